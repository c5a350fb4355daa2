"""Pinning tests of the CPU oracle (oracle/steel_oracle.cpp) against facts that do not come from its own code:
published inverses, libm, closed forms, brute force, definitions (first-fit coloring, KKT conditions),
invariants (momentum, energy).  The reference ships no physics fixtures or golden vectors, so these are
what pins the oracle."""
import ctypes
import math

import numpy as np
import pytest

from oracle_bind import OracleWorld, lib
from steel_webgpu_amd import SceneBuilder, pack_scene, workloads


def lowbias32_inverse(x):
    """Published inverse of lowbias32 (C. Wellons, hash-prospector, "lowbias32_r")."""
    M = 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x43021123) & M
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & M
    x ^= x >> 16
    return x


def test_hash_u32_inverts_with_published_inverse():
    L = lib()
    rng = np.random.default_rng(1)
    for x in list(rng.integers(0, 2**32, 2000, dtype=np.uint64)) + [0, 1, 0xFFFFFFFF, 0x80000000]:
        x = int(x)
        assert L.oracle_hash_u32(lowbias32_inverse(x)) == x
    assert L.oracle_hash_u32(0) == 0


def test_sin_cos_atan2_within_two_ulp_of_libm():
    L = lib()
    xs = np.concatenate([np.linspace(-20, 20, 4001), np.linspace(-1e-3, 1e-3, 101)]).astype(np.float32)
    for fn, ref in ((L.oracle_sinf, np.sin), (L.oracle_cosf, np.cos)):
        got = np.array([fn(float(x)) for x in xs], np.float32)
        want = ref(xs.astype(np.float64))
        ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - want)
        assert np.all(err <= 2 * ulp + 1e-9), float((err / ulp).max())
    ys = np.linspace(1e-5, 3, 301).astype(np.float32)
    for x in (-2.0, -0.5, 1e-3, 0.7, 3.0):
        got = np.array([L.oracle_atan2f(float(y), x) for y in ys])
        want = np.arctan2(ys.astype(np.float64), x)
        assert np.all(np.abs(got - want) <= 3 * np.spacing(np.abs(want).astype(np.float32)) + 1e-9)


def test_friction_pair_solve_meets_kkt_of_the_disk_problem():
    """min 1/2 (l-lo)^T K (l-lo) + v^T (l-lo) s.t. |l| <= R (omega 1): stationarity with a radial multiplier."""
    L = lib()
    rng = np.random.default_rng(2)
    out = np.zeros(2, np.float32)
    from oracle_bind import P
    for _ in range(400):
        A = rng.normal(size=(2, 2))
        K = A @ A.T + 0.05 * np.eye(2)
        v = rng.normal(size=2) * rng.choice([0.01, 1, 10])
        lo = rng.normal(size=2) * 0.3
        R = abs(rng.normal()) + 0.05
        if np.linalg.norm(lo) > R:
            lo = lo * (R / np.linalg.norm(lo)) * 0.99
        L.oracle_friction_pair_solve(*(float(x) for x in (K[0, 0], K[0, 1], K[1, 1], v[0], v[1], lo[0], lo[1], R, 1.0)), P(out))
        lam = out.astype(np.float64)
        g = K @ (lam - lo) + v  # gradient; KKT: g = -mu * lam, mu >= 0, or g = 0 inside
        nrm = np.linalg.norm(lam)
        assert nrm <= R * (1 + 1e-5)
        scale = np.linalg.norm(K) * max(np.linalg.norm(lam - lo), 1e-3) + np.linalg.norm(v)
        if nrm < R * (1 - 1e-4):
            assert np.linalg.norm(g) <= 2e-4 * scale
        else:
            # g antiparallel to lam
            cross = abs(g[0] * lam[1] - g[1] * lam[0])
            assert cross <= 2e-4 * scale * nrm
            assert g @ lam <= 2e-4 * scale * nrm


def _one_substep(scene, debug=True, f64=False):
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=f64)
    o.substeps(1, debug=debug)
    return pw, o


def _contacts(o):
    n = int(o.dbg_bufs["ncontacts"][0])
    return o.dbg_bufs["contacts"][:n * 24].reshape(n, 24)


def test_broadphase_equals_brute_force_all_pairs():
    scene = workloads.boxes64()
    pw, o = _one_substep(scene)
    NS = pw.NS
    aabb = o.poses[pw.h("W_P_AABB") * 4:(pw.h("W_P_AABB") + 2 * NS) * 4].reshape(NS, 2, 4)
    body = [int(pw.statF[(pw.h("W_SF_SHAPE") + s * 7) * 4]) for s in range(NS)]
    want = set()
    for i in range(NS):
        for j in range(i + 1, NS):
            if body[i] == body[j] or (body[i] < 0 and body[j] < 0):
                continue
            lo_i, hi_i, lo_j, hi_j = aabb[i, 0, :3], aabb[i, 1, :3], aabb[j, 0, :3], aabb[j, 1, :3]
            if np.any(lo_i > hi_j) or np.any(lo_j > hi_i):
                continue
            want.add((i, j))
    n = int(o.dbg_bufs["npairs"][0])
    got = set(map(tuple, o.dbg_bufs["pairs"][:2 * n].reshape(-1, 2).tolist()))
    assert got == want and len(want) > 0


def test_box_on_plane_four_corner_contacts():
    """A box resting axis-aligned at height z over the plane: 4 bottom corners, phi = z - hz, n = +z."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane()
    h = (0.1, 0.15, 0.05)
    z = 0.08
    body = b.addBody(xform=[0.3, -0.2, z, 0, 0, 0, 1])
    b.addShapeBox(body, hx=h[0], hy=h[1], hz=h[2])
    pw, o = _one_substep(b.finalize(solver=dict(substeps=1)))
    C = _contacts(o)
    assert len(C) == 4
    phi = C[:, 3]
    np.testing.assert_allclose(phi, z - h[2], atol=2e-6)
    np.testing.assert_allclose(C[:, 4:7], np.tile([0, 0, 1], (4, 1)), atol=1e-6)
    xs = sorted(map(tuple, np.round(C[:, :2], 5).tolist()))
    want = sorted((0.3 + sx * h[0], -0.2 + sy * h[1]) for sx in (-1, 1) for sy in (-1, 1))
    np.testing.assert_allclose(np.array(xs), np.array(want), atol=2e-5)
    # point midway between the surfaces
    np.testing.assert_allclose(C[:, 2], 0.5 * (z - h[2]), atol=2e-6)


def test_sphere_sphere_closest_points():
    b = SceneBuilder(fps=240)
    p1, p2, r1, r2 = np.array([0.0, 0.0, 1.0]), np.array([0.05, 0.12, 1.09]), 0.07, 0.05
    for p, r in ((p1, r1), (p2, r2)):
        bd = b.addBody(xform=list(p) + [0, 0, 0, 1], gravity=False)
        b.addShapeSphere(bd, radius=r)
    pw, o = _one_substep(b.finalize(solver=dict(substeps=1)))
    C = _contacts(o)
    assert len(C) == 1
    d = np.linalg.norm(p2 - p1)
    n = (p2 - p1) / d
    np.testing.assert_allclose(C[0, 3], d - r1 - r2, atol=2e-6)
    np.testing.assert_allclose(C[0, 4:7], n, atol=2e-6)
    np.testing.assert_allclose(C[0, :3], p1 + n * (r1 + 0.5 * (d - r1 - r2)), atol=2e-6)


def test_stacked_boxes_face_manifold_on_overlap():
    """A small box centred on a larger one: four contacts at the small box's bottom corners."""
    b = SceneBuilder(fps=240)
    lower = b.addBody(xform=[0, 0, 1.0, 0, 0, 0, 1], gravity=False)
    b.addShapeBox(lower, hx=0.3, hy=0.3, hz=0.1)
    gap = 0.01
    upper = b.addBody(xform=[0.02, -0.03, 1.0 + 0.1 + 0.05 + gap, 0, 0, 0, 1], gravity=False)
    b.addShapeBox(upper, hx=0.1, hy=0.08, hz=0.05)
    pw, o = _one_substep(b.finalize(solver=dict(substeps=1)))
    C = _contacts(o)
    assert len(C) == 4
    np.testing.assert_allclose(C[:, 3], gap, atol=3e-6)
    np.testing.assert_allclose(np.abs(C[:, 6]), 1.0, atol=1e-6)
    xs = sorted(map(tuple, np.round(C[:, :2], 5).tolist()))
    want = sorted((0.02 + sx * 0.1, -0.03 + sy * 0.08) for sx in (-1, 1) for sy in (-1, 1))
    np.testing.assert_allclose(np.array(xs), np.array(want), atol=3e-5)


def _sphere_points(n, r, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return (r * v / np.linalg.norm(v, axis=1, keepdims=True)).tolist()


def test_large_hulls_take_mpr_and_report_the_penetration_along_the_normal():
    """Hulls of 40 vertices take the MPR path.  Touching (the spheres they are sampled from overlap), the
    deepest contact's depth equals the brute-force separation of the two vertex sets along the contact normal
    (min over B minus max over A), and the normal roughly follows the line between the centers; 3 cm apart
    with 5 mm gaps they give no contact."""
    for seed, u, overlap in ((1, (0.1, -0.2, 1.0), 0.03), (2, (0.7, 0.3, 0.4), 0.012)):
        u = np.array(u) / np.linalg.norm(u)
        A, B = np.array(_sphere_points(40, 0.12, seed)), np.array(_sphere_points(40, 0.09, seed + 10))
        cA = np.array([0.0, 0.0, 1.0])
        for dist, expect in ((0.12 + 0.09 - overlap, True), (0.12 + 0.09 + 0.03, False)):
            cB = cA + u * dist
            b = SceneBuilder(fps=240)
            for c, P in ((cA, A), (cB, B)):
                bd = b.addBody(xform=list(c) + [0, 0, 0, 1], gravity=False)
                b.addShapeConvexHull(bd, P.tolist(), gap=0.005)
            scene = b.finalize(solver=dict(substeps=1))
            assert all(sh["hull"]["nv"] == 40 for sh in scene.json["shapes"] if sh.get("hull"))
            pw, o = _one_substep(scene, f64=True)
            C = _contacts(o)
            if not expect:
                assert len(C) == 0
                continue
            assert len(C) >= 1
            n = C[0, 4:7].astype(np.float64)
            assert np.all(np.abs(C[:, 4:7] - n) < 1e-6)
            assert n @ u > np.cos(np.radians(35))  # a face normal of the coarse hulls may win
            sep = ((cB + B) @ n).min() - ((cA + A) @ n).max()
            assert abs(C[:, 3].min() - sep) < 1e-5


def _rounded_box(h):
    """A box of half size h whose 12 edge midpoints stand 1 % proud (20 hull vertices: the MPR path)."""
    pts = [[sx * h, sy * h, sz * h] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    for ax in range(3):
        o = [k for k in range(3) if k != ax]
        for s1 in (-1, 1):
            for s2 in (-1, 1):
                p = [0.0, 0.0, 0.0]
                p[o[0]], p[o[1]] = s1 * h * 1.01, s2 * h * 1.01
                pts.append(p)
    return np.array(pts)


def test_large_hulls_crossed_edges_contact_along_the_edge_normal():
    """Two such boxes, one turned 45 degrees about x (an edge along x on top), the other 45 degrees about y (an
    edge along y below), stacked so the edges cross: one contact whose normal is the cross product of the edge
    directions (+z) and whose depth is minus the overlap of the edges (2 mm in, then 3 mm apart).  Face axes
    alone would report a tilted normal and a deep penetration."""
    from scipy.spatial.transform import Rotation as Rot
    P_ = _rounded_box(0.1)
    qa, qb = Rot.from_euler("x", 45, degrees=True), Rot.from_euler("y", 45, degrees=True)
    zA, zB = qa.apply(P_)[:, 2].max(), qb.apply(P_)[:, 2].min()
    for overlap in (0.002, -0.003):
        b = SceneBuilder(fps=240)
        a = b.addBody(xform=[0, 0, 1, *qa.as_quat()], gravity=False)
        b.addShapeConvexHull(a, P_.tolist(), gap=0.005)
        c = b.addBody(xform=[0, 0, 1 + zA - zB - overlap, *qb.as_quat()], gravity=False)
        b.addShapeConvexHull(c, P_.tolist(), gap=0.005)
        pw, o = _one_substep(b.finalize(solver=dict(substeps=1)), f64=True)
        C = _contacts(o)
        assert len(C) == 1
        np.testing.assert_allclose(C[0, 4:7], [0, 0, 1], atol=1e-6)
        assert abs(C[0, 3] + overlap) < 2e-7


def test_mpr_depth_is_the_minkowski_support_along_its_normal():
    """MPR (XenoCollide) on 40-vertex hulls under random relative poses: when the inflated shapes overlap it
    returns a unit normal n of the Minkowski difference D = B - A whose depth is D's support along n plus the
    inflation (brute force over the vertex sets), and no less than the true penetration depth, the minimum of
    that support over 20000 directions; far apart it reports a miss."""
    from scipy.spatial.transform import Rotation as Rot
    from oracle_bind import mpr
    rng = np.random.default_rng(7)
    b = SceneBuilder(fps=240)
    A, B = np.array(_sphere_points(40, 0.12, 3)), np.array(_sphere_points(40, 0.08, 4))
    for P_ in (A, B):
        bd = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1], gravity=False)
        b.addShapeConvexHull(bd, P_.tolist())
    pw = pack_scene(b.finalize(solver=dict(substeps=1)))
    so = pw.h("W_SF_SHAPE")
    ha, hb = (int(pw.statF[(so + s * 7) * 4 + 1]) for s in (0, 1))
    dirs = rng.normal(size=(20000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    hits = 0
    for trial in range(12):
        rot = Rot.random(random_state=trial)
        q = rot.as_quat()
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        t = u * rng.uniform(0.1, 0.22)
        r = 0.005
        n_d = mpr(pw, ha, hb, q, t, r)
        Bw = rot.apply(B) + t
        hD = lambda n: (Bw @ n).max() - (A @ n).min()
        pen = min((Bw @ dirs.T).max(axis=0) - (A @ dirs.T).min(axis=0)) + 2 * r
        if n_d[3] < 0:
            assert pen > 0, (trial, pen)  # reported misses are real misses of the inflated shapes
            continue
        hits += 1
        n = n_d[:3]
        assert abs(np.linalg.norm(n) - 1) < 1e-9
        assert abs(n_d[3] - (hD(n) + 2 * r)) < 2e-6
        assert n_d[3] >= -pen - 1e-9
    assert hits >= 4
    assert mpr(pw, ha, hb, (0, 0, 0, 1), (1.0, 0.2, 0), 0.005)[3] < 0


def test_cylinder_support_is_the_extreme_point_of_the_solid_cylinder():
    """The analytic cylinder support against a dense sampling of the cap rims (3600 points each): no sample
    reaches further along d, and the point lies on the cylinder."""
    from oracle_bind import gsup
    r, hh = 0.05, 0.1
    b = SceneBuilder(fps=240)
    bd = b.addBody(xform=[0, 0, 1, 0, 0, 0, 1], gravity=False)
    b.addShapeCylinder(bd, radius=r, halfHeight=hh)
    pw = pack_scene(b.finalize(solver=dict(substeps=1)))
    r, hh = _f32(r), _f32(hh)  # the packed dimensions
    a = np.linspace(0, 2 * np.pi, 3600, endpoint=False)
    rim = np.concatenate([np.stack([r * np.cos(a), r * np.sin(a), np.full_like(a, z)], 1) for z in (-hh, hh)])
    rng = np.random.default_rng(1)
    for d in list(rng.normal(size=(50, 3))) + [np.array([0.3, -0.4, 0.0]), np.array([0, 0, -1.0])]:
        p = gsup(pw, 0, d)
        assert p @ d >= (rim @ d).max() - 1e-9
        assert np.hypot(p[0], p[1]) <= r * (1 + 1e-9) and abs(p[2]) <= hh * (1 + 1e-9)
        if d[2] == 0:
            assert p[2] == 0


def _cylinder_on(ground, quat, z, r=0.05, hh=0.1):
    b = SceneBuilder(fps=240)
    if ground == "plane":
        b.addGroundPlane(gap=0.001)
    else:
        bd = b.addBody(xform=[0, 0, -0.2, 0, 0, 0, 1], gravity=False)
        b.addShapeBox(bd, hx=0.5, hy=0.5, hz=0.2, gap=0.001)
    bd = b.addBody(xform=[0.1, -0.05, z, *quat], gravity=False)
    b.addShapeCylinder(bd, radius=r, halfHeight=hh, gap=0.001)
    return b.finalize(solver=dict(substeps=1))


def test_upright_cylinder_on_plane_touches_at_three_rim_points():
    """Upright, sunk by 2 mm: the deepest rim point and two more at 120 degrees on the bottom rim, each at depth
    -2 mm, midway between the surfaces, normal +z."""
    r, hh, dz = 0.05, 0.1, 0.002
    pw, o = _one_substep(_cylinder_on("plane", (0, 0, 0, 1), hh - dz), f64=True)
    C = _contacts(o)
    assert len(C) == 3
    np.testing.assert_allclose(C[:, 3], -dz, atol=1e-6)
    np.testing.assert_allclose(C[:, 4:7], np.tile([0, 0, 1], (3, 1)), atol=1e-7)
    rel = C[:, :2] - [0.1, -0.05]
    np.testing.assert_allclose(np.hypot(rel[:, 0], rel[:, 1]), r, atol=1e-6)
    ang = np.sort(np.degrees(np.arctan2(rel[:, 1], rel[:, 0])) % 360)
    np.testing.assert_allclose(np.diff(ang), 120.0, atol=1e-3)
    np.testing.assert_allclose(C[:, 2], -dz / 2, atol=1e-7)


def test_cylinder_lying_on_plane_touches_at_both_ends_of_its_bottom_line():
    """Lying along y, sunk by 2 mm: its two bottom end points (y = +-hh) at depth -2 mm."""
    r, hh, dz = 0.05, 0.1, 0.002
    s = np.sqrt(0.5)
    pw, o = _one_substep(_cylinder_on("plane", (s, 0, 0, s), r - dz), f64=True)
    C = _contacts(o)
    assert len(C) == 2
    np.testing.assert_allclose(C[:, 3], -dz, atol=1e-6)
    np.testing.assert_allclose(np.sort(C[:, 1]), [-0.05 - hh, -0.05 + hh], atol=1e-6)
    np.testing.assert_allclose(C[:, 0], 0.1, atol=1e-6)


def test_cylinder_rolling_on_a_box_keeps_its_contacts_under_the_axis():
    """A cylinder lying on a box (a discrete shape; 48-vertex hull, MPR path) within 2 degrees of flat rolls: its
    contact points are moved onto the plane through its axis and the normal, so they sit straight under the axis,
    at depth -2 mm along +z."""
    r, hh, dz = 0.05, 0.1, 0.002
    s = np.sqrt(0.5)
    pw, o = _one_substep(_cylinder_on("box", (s, 0, 0, s), r - dz), f64=True)
    C = _contacts(o)
    assert len(C) >= 2
    np.testing.assert_allclose(C[:, 0], 0.1, atol=1e-6)
    np.testing.assert_allclose(np.abs(C[:, 6]), 1.0, atol=1e-6)
    assert abs(C[:, 3].min() + dz) < 2e-5
    assert C[:, 1].min() < -0.05 - 0.5 * hh and C[:, 1].max() > -0.05 + 0.5 * hh


def test_contacts_sorted_by_canonical_key():
    pw, o = _one_substep(workloads.boxes64())
    C = _contacts(o)
    pair = (C[:, 8].astype(np.int64) + 1) | ((C[:, 9].astype(np.int64) + 1) << 16)
    feat = C[:, 10].astype(np.int64) | (C[:, 11].astype(np.int64) << 16)
    key = (pair << 32) | feat
    assert np.all(np.diff(key) > 0)


def test_coloring_is_first_fit_in_walk_order():
    """Each row takes the lowest color no earlier row of either of its bodies holds (the world never
    conflicts); colors past 63 fold into the serial overflow color 64.  Rows are in walk order."""
    for scene in (workloads.boxes64(), workloads.quadruped_world(3)):
        pw = pack_scene(scene)
        o = OracleWorld(pw)
        o.substeps(600, debug=True)
        nr = int(o.dbg_bufs["nrows"][0])
        info = o.dbg_bufs["rowinfo"][:nr * 4].reshape(nr, 4)
        used = {}
        ncol = 64
        for r in range(nr):
            key = int(info[r, 3])
            a, b = (key & 0xFFFF) - 1, (key >> 16) - 1
            busy = used.get(a, set()) | used.get(b, set())
            c = next((k for k in range(ncol) if k not in busy), ncol)
            assert info[r, 1] == c
            if c < ncol:
                for x in (a, b):
                    if x >= 0:
                        used.setdefault(x, set()).add(c)
        assert nr > 10


def _free_box_scene(z=5.0, gravity=True, twist=(0, 0, 0, 0, 0, 0)):
    b = SceneBuilder(fps=240)
    bd = b.addBody(xform=[0.1, 0.2, z, 0, 0, 0, 1], gravity=gravity, twist=twist)
    b.addShapeBox(bd, hx=0.1, hy=0.2, hz=0.3)
    return b.finalize(solver=dict(substeps=4))


def test_ballistic_free_fall_closed_form_fp64():
    pw = pack_scene(_free_box_scene())
    o = OracleWorld(pw, f64=True)
    n = 40
    o.substeps(n)
    dt = pw.hf("W_DT")
    g = -9.81
    gz = float(np.float32(g))
    assert o.vel[2] == pytest.approx(n * dt * gz, rel=1e-12)
    zexp = 5.0 + dt * dt * gz * n * (n + 1) / 2
    assert o.poses[2] == pytest.approx(zexp, rel=1e-12)
    assert o.poses[0] == float(np.float32(0.1)) and o.poses[1] == float(np.float32(0.2))


def test_gyroscopic_step_conserves_kinetic_energy():
    """Free asymmetric box spinning with no damping and no gravity: Cayley microsteps keep w^T I w."""
    scene = _free_box_scene(gravity=False, twist=(0, 0, 0, 3.0, -2.0, 5.0))
    scene.json["bodies"]["angular_damping"] = [0.0]
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=True)
    I = np.array(scene.json["bodies"]["inertia"][0], np.float32).astype(np.float64).reshape(3, 3)

    def energy():
        q = o.poses[4:8]
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        wv = o.vel[4:7]
        return wv @ (R @ I @ R.T) @ wv

    e0 = energy()
    o.substeps(200)
    # explicit Euler on the same spin drifts ~1e-2 here; the Cayley microsteps stay within the
    # fixed-point tolerance of each microstep
    assert energy() == pytest.approx(e0, rel=1e-6)


def test_collision_conserves_linear_momentum():
    b = SceneBuilder(fps=240)
    b1 = b.addBody(xform=[0, 0, 1, 0, 0, 0, 1], gravity=False, twist=(1.0, 0.2, 0, 0, 0, 0))
    b.addShapeBox(b1, hx=0.1, hy=0.1, hz=0.1, mu=0.5)
    b2 = b.addBody(xform=[0.25, 0.05, 1.02, 0, 0, 0, 1], gravity=False, twist=(-1.5, 0, 0.1, 0, 0, 0))
    b.addShapeBox(b2, hx=0.08, hy=0.12, hz=0.1, mu=0.5)
    scene = b.finalize(solver=dict(substeps=4))
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=True)
    inv_m = np.array([pw.statF[(pw.h("W_SF_BODY") + b * 8) * 4] for b in range(pw.NB)], np.float64)
    m = 1.0 / inv_m  # the masses the step integrates with

    def mom():
        v = o.vel.reshape(-1, 8)[:, :3]
        return (m[:, None] * v).sum(0)

    p0 = mom()
    o.substeps(120)
    assert o.stats[1] > 0  # they touched
    np.testing.assert_allclose(mom(), p0, atol=1e-9 * np.abs(p0).max())


def test_floating_base_free_fall_has_zero_joint_acceleration():
    """Articulated-body free velocity (qd_hat) of an untouched floating quadruped at rest: the whole tree falls,
    base acceleration g, joint accelerations 0 (exact for any configuration)."""
    scene = workloads.quadruped_world(7)
    for p in scene.json["bodies"]["q"]:
        p[2] += 3.0
    for j, t in enumerate(scene.json["joints"]["type"]):
        if t == 4:
            scene.json["joints"]["q"][scene.json["joints"]["q_start"][j] + 2] += 3.0
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    A = pw.A
    dt = pw.hf("W_DT")
    qdh = o.art[A["A_QDH"]:A["A_QDH"] + pw.ND]
    g = np.array([0, 0, np.float32(-9.81)], np.float64)
    np.testing.assert_allclose(qdh[0:3], dt * g, atol=1e-12)
    np.testing.assert_allclose(qdh[3:], 0.0, atol=1e-12)


def _rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _revolute_chain(rng, n, dof=None):
    """Fixed-base chain of n capsule links on revolute joints with random unit axes and angles."""
    b = SceneBuilder(fps=240)
    parent, prev = -1, 0.0
    for i in range(n):
        l = float(rng.uniform(0.1, 0.3))
        bd = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1])
        b.addShapeCapsule(bd, radius=0.03, halfHeight=l / 2, xform=[0, 0, l / 2, 0, 0, 0, 1])
        ax = rng.normal(size=3)
        b.addJointRevolute(parent, bd, axis=list(ax / np.linalg.norm(ax)), parentXform=[0, 0, prev, 0, 0, 0, 1], q=float(rng.uniform(-1, 1)),
                           **(dof(i) if dof else {}))
        parent, prev = bd, l
    return b.finalize(solver=dict(substeps=1))


def _crba(scene):
    """Joint-space mass matrix H and gravity torque of a fixed-base revolute chain, from the composite rigid-body
    formula (sum over links of COM Jacobians) with forward kinematics computed here from the scene's joint arrays
    (body frame = joint child frame; X_c identity, X_p a translation)."""
    from scipy.spatial.transform import Rotation as Rot
    S, J = scene.json, scene.json["joints"]
    n = J["count"]
    R_prev, p_prev = np.eye(3), np.zeros(3)
    axes_w, orig_w, coms, Rs = [], [], [], []
    for j in range(n):
        xp = np.array(J["X_p"][j][:3])
        a = np.array(J["axis"][J["qd_start"][j]])
        o = p_prev + R_prev @ xp
        aw = R_prev @ a
        R = R_prev @ Rot.from_rotvec(a / np.linalg.norm(a) * J["q"][J["q_start"][j]]).as_matrix()
        axes_w.append(aw); orig_w.append(o); Rs.append(R)
        coms.append(o + R @ np.array(S["bodies"]["com"][J["child"][j]]))
        R_prev, p_prev = R, o
    H = np.zeros((n, n))
    tau = np.zeros(n)
    g = np.array([0, 0, float(np.float32(-9.81))])
    for l in range(n):
        bd = J["child"][l]
        m = S["bodies"]["mass"][bd]
        Iw = Rs[l] @ np.array(S["bodies"]["inertia"][bd]).reshape(3, 3) @ Rs[l].T
        Jv, Jw = np.zeros((3, n)), np.zeros((3, n))
        for k in range(l + 1):
            Jw[:, k] = axes_w[k]
            Jv[:, k] = np.cross(axes_w[k], coms[l] - orig_w[k])
        H += m * Jv.T @ Jv + Jw.T @ Iw @ Jw
        tau += Jv.T @ (m * g)
    return H, tau


def test_fixed_base_chain_free_acceleration_matches_crba():
    """qd_hat of a fixed-base revolute chain at rest = q_dot + dt * H^-1 tau_gravity."""
    scene = _revolute_chain(np.random.default_rng(5), 7)
    H, tau = _crba(scene)
    qdd = np.linalg.solve(H, tau)
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    dt = pw.hf("W_DT")
    qdh = o.art[pw.A["A_QDH"]:pw.A["A_QDH"] + len(qdd)]
    np.testing.assert_allclose(qdh / dt, qdd, rtol=2e-6)


def _f32(x):
    return float(np.float32(x))


def test_implicit_drives_springs_and_damping_match_backward_euler_joint_dynamics():
    """A chain at rest with a PD drive, armature, passive damping and a passive spring on every joint.  Backward
    Euler on the drive and damping (forces evaluated at q + dt qd', qd'), forward on the spring, gives
    (H + A + dt C + dt^2 K) qdd = tau_g + K (q_tgt - q) + C_d qd_tgt + K_s (q_ref - q), A armature, K drive
    stiffness, C = drive + passive damping; qd_hat = dt qdd at rest.  Unsaturated: efforts are unlimited."""
    rng = np.random.default_rng(11)
    n = 6
    cfg = [dict(targetKe=_f32(rng.uniform(20, 400)), targetKd=_f32(rng.uniform(0.5, 20)), targetPos=_f32(rng.uniform(-1, 1)),
                targetVel=_f32(rng.uniform(-2, 2)), armature=_f32(rng.uniform(0, 0.02)), damping=_f32(rng.uniform(0, 0.5)), effortLimit=0.0)
           for _ in range(n)]
    scene = _revolute_chain(rng, n, dof=lambda i: cfg[i])
    ks = [_f32(rng.uniform(0, 30)) for _ in range(n)]
    qref = [_f32(rng.uniform(-1, 1)) for _ in range(n)]
    scene.json["joints"]["spring_stiffness"] = ks
    scene.json["joints"]["spring_ref"] = qref
    H, tau = _crba(scene)
    pw = pack_scene(scene)
    dt = pw.hf("W_DT")
    q = np.array(scene.json["joints"]["q"], np.float64)
    K = np.array([c["targetKe"] for c in cfg])
    Cd = np.array([c["targetKd"] for c in cfg])
    C = Cd + np.array([c["damping"] for c in cfg])
    A = np.array([c["armature"] for c in cfg])
    tq = np.array([c["targetPos"] for c in cfg], np.float64)
    tv = np.array([c["targetVel"] for c in cfg], np.float64)
    rhs = tau + K * (tq - q) + Cd * tv + np.array(ks) * (np.array(qref) - q)
    qdd = np.linalg.solve(H + np.diag(A + dt * C + dt * dt * K), rhs)
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    qdh = o.art[pw.A["A_QDH"]:pw.A["A_QDH"] + n]
    np.testing.assert_allclose(qdh / dt, qdd, rtol=2e-6)


def _vertical_hinge(**dof):
    """One capsule on a vertical revolute axis: gravity exerts no torque about it and the joint-space equation
    is I qdd = tau exactly, I the link's inertia about the axis."""
    b = SceneBuilder(fps=240)
    bd = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1])
    b.addShapeCapsule(bd, radius=0.03, halfHeight=0.1, xform=[0.1, 0, 0, 0, 0.7071068, 0, 0.7071068])
    b.addJointRevolute(-1, bd, axis=[0, 0, 1], parentXform=[0, 0, 1, 0, 0, 0, 1], q=0.2, **dof)
    scene = b.finalize(solver=dict(substeps=1))
    scene.json["bodies"]["angular_damping"] = [0.0]
    return scene


def test_moving_joint_drive_velocity_update_is_backward_euler():
    """Moving hinge (qd != 0): qd' solves I (qd' - qd) / dt = K (q_tgt - q - dt qd') + C_d (qd_tgt - qd') - c qd'
    + K_s (q_ref - q)."""
    ke, kd, damp, arm, ks, qref, tq, tv, qd = 150.0, 4.0, 0.3, 0.01, 12.0, -0.4, 0.9, -1.5, 2.5
    scene = _vertical_hinge(targetKe=ke, targetKd=kd, targetPos=tq, targetVel=tv, damping=damp, armature=arm, effortLimit=0.0)
    scene.json["joints"]["spring_stiffness"] = [ks]
    scene.json["joints"]["spring_ref"] = [qref]
    H, _ = _crba(scene)
    pw = pack_scene(scene)
    pw.art0[pw.A["A_QD"]] = qd
    dt = pw.hf("W_DT")
    q = _f32(0.2)
    qdv = _f32(qd)
    I = H[0, 0] + _f32(arm)
    qd_new = (I * qdv + dt * (_f32(ke) * (_f32(tq) - q) + _f32(kd) * _f32(tv) + _f32(ks) * (_f32(qref) - q))) / (
        I + dt * (_f32(kd) + _f32(damp)) + dt * dt * _f32(ke))
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    np.testing.assert_allclose(o.art[pw.A["A_QDH"]], qd_new, rtol=2e-6)


def test_saturated_drive_applies_exactly_the_effort_limit():
    """A stiff drive far from its target on a light hinge saturates: the joint then receives +-effort as an
    explicit torque (no implicit stiffness), qd_hat = dt effort / (I + armature) from rest."""
    effort = 0.05
    scene = _vertical_hinge(targetKe=1e4, targetKd=50.0, targetPos=1.5, effortLimit=effort, armature=0.002)
    H, _ = _crba(scene)
    pw = pack_scene(scene)
    dt = pw.hf("W_DT")
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    np.testing.assert_allclose(o.art[pw.A["A_QDH"]], dt * _f32(effort) / (H[0, 0] + _f32(0.002)), rtol=2e-6)
    # toward a target below the angle the torque is -effort
    scene = _vertical_hinge(targetKe=1e4, targetKd=50.0, targetPos=-1.5, effortLimit=effort, armature=0.002)
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    np.testing.assert_allclose(o.art[pw.A["A_QDH"]], -dt * _f32(effort) / (H[0, 0] + _f32(0.002)), rtol=2e-6)


def _spin_hinge(qd, vlim, frac=None, **dof):
    scene = _vertical_hinge(velocityLimit=vlim, **dof)
    scene.json["solver"]["enable_joint_velocity_limits"] = True
    if frac is not None:
        scene.json["solver"]["velocity_limit_activation_fraction"] = frac
    pw = pack_scene(scene)
    pw.art0[pw.A["A_QD"]] = qd
    return pw


def test_velocity_limit_caps_a_spinning_hinge_and_leaves_slower_ones_alone():
    """A hinge spinning at 5 rad/s under a 2 rad/s limit leaves its first step at the limit (nothing else acts
    on it: vertical axis, no damping) and moves by dt times the limit; at 1 rad/s with the rows active only
    beyond 0.9 of the limit nothing changes."""
    pw = _spin_hinge(5.0, 2.0)
    dt = pw.hf("W_DT")
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    A = pw.A
    assert abs(o.art[A["A_QD"]] - 2.0) <= 2e-7
    assert abs(o.art[A["A_Q"]] - (_f32(0.2) + dt * 2.0)) <= 1e-7
    pw = _spin_hinge(1.0, 2.0, frac=0.9)
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    assert o.art[A["A_VLG"]] == 0.0 and o.art[A["A_AVL"]] == 0.0
    assert o.art[A["A_QD"]] == 1.0
    # the same hinge without the switch keeps spinning at 5 rad/s
    scene = _vertical_hinge(velocityLimit=2.0)
    pw = pack_scene(scene)
    pw.art0[pw.A["A_QD"]] = 5.0
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    assert o.art[A["A_QD"]] == 5.0


def test_velocity_limit_holds_against_a_drive():
    """A stiff drive toward a far target accelerates the hinge; the velocity limit holds |qd| at the limit
    while the drive would go faster, and the joint still arrives."""
    pw = _spin_hinge(0.0, 1.5, targetKe=400.0, targetKd=1.0, targetPos=2.0, effortLimit=0.0)
    o = OracleWorld(pw)
    A = pw.A
    peak = 0.0
    for _ in range(480):
        o.substeps(1)
        peak = max(peak, abs(float(o.art[A["A_QD"]])))
    assert 1.5 * (1 - 1e-3) <= peak <= 1.5 * (1 + 1e-6)
    assert abs(o.art[A["A_Q"]] - 2.0) < 0.05


def test_velocity_limit_prescale_scales_every_scalar_dof_of_the_articulation():
    """The pre-solve prescale multiplies all scalar dof velocities of an articulation by the ratio that meets
    its most exceeded limit (wgsl.js:1828-1845).  A 3-link chain with one dof at 4x its 0.5 rad/s limit
    (ratio 1/4, exact in binary) has the same free velocity qd_hat as the same chain started at qd / 4 without
    velocity limits, and a different one from the chain started at qd."""
    qd = np.array([0.3, 2.0, -0.7], np.float32)

    def qd_hat(limits, start):
        scene = _revolute_chain(np.random.default_rng(2), 3, dof=lambda i: dict(velocityLimit=0.5 if i == 1 else 1e6))
        scene.json["solver"]["enable_joint_velocity_limits"] = limits
        pw = pack_scene(scene)
        pw.art0[pw.A["A_QD"]:pw.A["A_QD"] + 3] = start
        o = OracleWorld(pw, f64=True)
        o.substeps(1)
        return o.art[pw.A["A_QDH"]:pw.A["A_QDH"] + 3]

    limited = qd_hat(True, qd)
    assert np.array_equal(limited, qd_hat(False, qd * np.float32(0.25)))
    assert np.abs(limited - qd_hat(False, qd)).max() > 0.1


def _friction_hinge(qd, **dof):
    scene = _vertical_hinge(**dof)
    scene.json["solver"]["enable_joint_friction"] = True
    return scene


def test_joint_friction_removes_at_most_its_bound_per_step_and_stops_slow_joints():
    """Coulomb joint friction is an impulse bounded by friction * dt: a fast hinge loses exactly
    friction * dt / I of speed in a substep (nothing else acts about its vertical axis), a slow one stops."""
    fr = 0.02
    scene = _friction_hinge(0.0, friction=fr)
    H, _ = _crba(scene)
    pw = pack_scene(scene)
    dt = pw.hf("W_DT")
    pw.art0[pw.A["A_QD"]] = 3.0
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    np.testing.assert_allclose(o.art[pw.A["A_QD"]], 3.0 - dt * _f32(fr) / H[0, 0], rtol=1e-6)
    pw.art0[pw.A["A_QD"]] = 0.3 * dt * _f32(fr) / H[0, 0]
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    assert abs(o.art[pw.A["A_QD"]]) < 1e-6 * dt * fr / H[0, 0]


def test_load_dependent_joint_friction_bound_is_coefficient_times_transmitted_weight():
    """With a load coefficient c the bound grows by c times the force the joint transmits (force plus torque
    magnitude of the articulated force at the child COM).  A vertical hinge spinning at w carries the link's
    weight and its centripetal force, m sqrt(g^2 + (w^2 r)^2) for a COM at r from the axis, and no torque
    about the COM (w is along a principal axis), so it loses c m sqrt(g^2 + (w^2 r)^2) dt / I of speed."""
    c = 0.01
    scene = _friction_hinge(0.0)
    scene.json["joints"]["friction_load_coefficient"] = [c]
    H, _ = _crba(scene)
    pw = pack_scene(scene)
    assert pw.h("W_JFL") == 1
    dt = pw.hf("W_DT")
    m = scene.json["bodies"]["mass"][0]
    pw.art0[pw.A["A_QD"]] = 3.0
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    r = float(np.hypot(*scene.json["bodies"]["com"][0][:2]))
    fm = m * np.hypot(float(np.float32(9.81)), 9.0 * r)
    np.testing.assert_allclose(o.art[pw.A["A_QD"]], 3.0 - dt * _f32(c) * fm / H[0, 0], rtol=1e-6)


def _flying_box(twist, quat=(0, 0, 0, 1), vmax=None, wmax=None):
    b = SceneBuilder(fps=240)
    bd = b.addBody(xform=[0, 0, 5, *quat], twist=twist)
    b.addShapeBox(bd, hx=0.2, hy=0.1, hz=0.05)
    scene = b.finalize(solver=dict(substeps=1))
    B = scene.json["bodies"]
    B["angular_damping"] = [0.0]
    B["max_linear_velocity"] = [vmax]
    B["max_angular_velocity"] = [wmax]
    return scene


def test_free_body_velocity_limits_clamp_v_and_push_w_about_the_exceeded_axis():
    """Linear speed is boxed per axis; an angular component over its limit is brought back by an angular
    impulse about that world axis, so I * dw is parallel to the axis (a rotated box makes I non-diagonal)."""
    from scipy.spatial.transform import Rotation as Rot
    quat = Rot.from_euler("xyz", [0.4, -0.7, 1.1]).as_quat()
    scene = _flying_box([3.0, -0.2, 0.5, 0.3, -0.1, 3.0], quat, vmax=1.0, wmax=2.5)
    pw = pack_scene(scene)
    assert pw.h("W_VLIM_FREE") == 1
    o = OracleWorld(pw, f64=True)
    o.substeps(1)
    v, w = o.vel[0:3], o.vel[4:7]
    assert v[0] == 1.0 and abs(v[1] + 0.2) < 1e-6
    assert abs(w[2] - 2.5) < 1e-9 and np.abs(w).max() <= 2.5 + 1e-9
    # the same box without limits: nothing is clamped, and the difference is the push alone
    o2 = OracleWorld(pack_scene(_flying_box([3.0, -0.2, 0.5, 0.3, -0.1, 3.0], quat)), f64=True)
    o2.substeps(1)
    assert o2.vel[0] == 3.0
    R = Rot.from_quat(quat).as_matrix()
    I = R @ np.array(scene.json["bodies"]["inertia"][0]).reshape(3, 3) @ R.T
    L = I @ (w - o2.vel[4:7])
    assert abs(L[0]) < 1e-6 * abs(L[2]) and abs(L[1]) < 1e-6 * abs(L[2]) and L[2] < 0  # float32 packed inertia
    # a component below -limit is pushed up to it
    o3 = OracleWorld(pack_scene(_flying_box([0, 0, 0, 0.2, -3.0, 0.4], quat, wmax=2.5)), f64=True)
    o3.substeps(1)
    assert abs(o3.vel[5] + 2.5) < 1e-9


def _path_hinge(path, fps=60, substeps=4, **kw):
    scene = _vertical_hinge()
    scene.json["fps"] = fps
    scene.json["substeps"] = scene.json["solver"]["substeps"] = substeps
    return workloads.with_target_path(scene, np.asarray(path).reshape(-1, 1), **kw)


def test_recorded_path_substeps_average_to_each_frames_sample():
    """The path is sampled at each substep's midpoint half a frame back, so over the substeps of frame k the
    interpolated targets average to sample k (a linear path), and frame 0 starts clamped at sample 0; held
    paths give sample k at every substep of frame k."""
    path = 0.05 * np.arange(20)
    pw = pack_scene(_path_hinge(path))
    A = pw.A["A_TGT"]
    o = OracleWorld(pw, f64=True)
    tg = []
    for k in range(12 * 4):
        o.substeps(1)
        tg.append(o.art[A])
    tg = np.array(tg).reshape(12, 4)
    assert tg[0, 0] == np.float32(path[0])
    np.testing.assert_allclose(tg[1:].mean(axis=1), np.float32(path[1:12]), atol=2e-6)
    pw = pack_scene(_path_hinge(path, held=True))
    o = OracleWorld(pw, f64=True)
    for k in range(6):
        got = []
        for s_ in range(4):
            o.substeps(1)
            got.append(o.art[A])
        np.testing.assert_allclose(got, np.float32(path[k]), atol=1e-6)


def test_recorded_path_integral_term_and_user_override():
    """The integral term of the first dofs grows by 3 * frame_dt * the clamped error once per frame, saturating at
    0.12 (a hinge that does not move, target 1 rad away); once the user sets targets the path stops driving."""
    pw = pack_scene(_path_hinge(np.full(200, 1.0), integral=1))
    A = pw.A
    o = OracleWorld(pw, f64=True)
    o.frames(10)
    assert abs(o.art[A["A_INT"]] - 10 * 3 * _f32(1 / 60) * 0.08) < 1e-6
    np.testing.assert_allclose(o.art[A["A_TGT"]], 1.0 + o.art[A["A_INT"]], atol=1e-6)
    o.frames(30)
    assert abs(o.art[A["A_INT"]] - 0.12) < 1e-7
    pw = pack_scene(_path_hinge(np.full(200, 1.0)))
    o = OracleWorld(pw, f64=True)
    o.count[1] = 1
    before = o.art[A["A_TGT"]]
    o.frames(5)
    assert o.art[A["A_TGT"]] == before


def _sleepy(scene, quiet=0.25):
    scene.json["solver"].update(enable_sleeping=True, sleep_quiet_time=quiet)
    return pack_scene(scene)


def test_resting_stack_falls_asleep_together_and_freezes():
    """Two boxes stacked on the ground: supported and quiet, their ages grow by dt per substep; the island
    (joined by their contact) sleeps in the same substep once both reached the quiet time, after which they
    stay exactly still with zero velocity.  A box in free fall never sleeps (not supported)."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane(gap=0.002)
    lo = b.addBody(xform=[0, 0, 0.1, 0, 0, 0, 1])
    b.addShapeBox(lo, hx=0.2, hy=0.2, hz=0.1, gap=0.002)
    hi = b.addBody(xform=[0.02, 0, 0.25, 0, 0, 0, 1])
    b.addShapeBox(hi, hx=0.1, hy=0.1, hz=0.05, gap=0.002)
    fl = b.addBody(xform=[2, 0, 50, 0, 0, 0, 1])
    b.addShapeBox(fl, hx=0.1, hy=0.1, hz=0.1, gap=0.002)
    pw = _sleepy(b.finalize(solver=dict(substeps=4)))
    S, A = pw.A["A_SLP"], pw.A
    dt = pw.hf("W_DT")
    o = OracleWorld(pw)
    asleep_at = None
    for k in range(4 * 240):
        o.substeps(1)
        sl = o.art[S:S + 12].reshape(3, 4)
        assert sl[0, 1] == sl[1, 1]          # one island
        assert sl[2, 1] == 0 and sl[2, 0] == 0  # free fall: never supported
        if sl[0, 1] == 1 and asleep_at is None:
            asleep_at = k
            assert o.stats[8] == 2
            break
    assert asleep_at is not None
    # ages reached the quiet time, having grown by dt per substep while quiet
    assert np.all(sl[:2, 0] == np.float32(0.25))
    p0, v0 = o.poses[:16].copy(), o.vel[:16].copy()
    o.substeps(40)
    assert np.array_equal(o.poses[:16], p0) and np.all(o.vel[:16] == 0)


def test_age_grows_by_dt_per_quiet_substep():
    """A box at rest on the ground: its age is k * dt (float32 sums) after k quiet substeps, until the quiet time."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane(gap=0.002)
    bd = b.addBody(xform=[0, 0, 0.1, 0, 0, 0, 1])
    b.addShapeBox(bd, hx=0.2, hy=0.2, hz=0.1, gap=0.002)
    pw = _sleepy(b.finalize(solver=dict(substeps=4)), quiet=1.0)
    S = pw.A["A_SLP"]
    dt = np.float32(pw.hf("W_DT"))
    o = OracleWorld(pw)
    ages = []
    for k in range(60):
        o.substeps(1)
        ages.append(o.art[S])
    # the first substeps may still see the settling motion; from the first nonzero age on it grows by dt
    ages = np.array(ages, np.float32)
    k0 = int(np.argmax(ages > 0))
    acc = np.float32(0)
    for a in ages[k0:]:
        acc = np.float32(acc + dt)
        assert a == acc


def test_driven_articulation_never_sleeps():
    """Drives veto sleep (robots follow their targets), even at rest on a fixed base."""
    scene = _vertical_hinge(targetKe=10.0, targetKd=1.0, targetPos=0.2, effortLimit=0.0)
    pw = _sleepy(scene, quiet=0.05)
    o = OracleWorld(pw)
    o.frames(60)
    S = pw.A["A_SLP"]
    node = pw.NB  # the articulation's node
    assert o.art[S + node * 4 + 1] == 0 and o.art[S + node * 4] == 0


def test_mimic_joint_holds_its_linear_relation_while_the_leader_moves():
    """A follower hinge mimics its leader (q_f = c0 + c1 q_l, here c0 = 0.1, c1 = -0.5) through a dense bilateral
    row with Baumgarte feedback: starting off the relation it converges onto it, and it stays on it (and the
    velocities with it, qd_f = c1 qd_l) while a joint torque swings the leader through a third of a radian."""
    b = SceneBuilder(fps=240)
    l1 = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1])
    b.addShapeCapsule(l1, radius=0.03, halfHeight=0.1, xform=[0.1, 0, 0, 0, 0.7071068, 0, 0.7071068])
    j1 = b.addJointRevolute(-1, l1, axis=[0, 0, 1], parentXform=[0, 0, 1, 0, 0, 0, 1], q=0.0)
    l2 = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1])
    b.addShapeCapsule(l2, radius=0.03, halfHeight=0.1, xform=[0.1, 0, 0, 0, 0.7071068, 0, 0.7071068])
    j2 = b.addJointRevolute(l1, l2, axis=[0, 0, 1], parentXform=[0.2, 0, 0, 0, 0, 0, 1], q=0.3)
    scene = b.finalize(solver=dict(substeps=4))
    J = scene.json["joints"]
    J["mimic_joint"][j2] = j1
    J["mimic_coeffs"][j2] = [0.1, -0.5]
    B = scene.json["bodies"]
    B["angular_damping"] = [0.0, 0.0]
    pw = pack_scene(scene)
    assert pw.h("W_NMIMIC") == 1
    A = pw.A
    o = OracleWorld(pw)
    err = lambda: o.art[A["A_Q"] + 1] - (0.1 - 0.5 * o.art[A["A_Q"]])
    assert abs(err() - 0.2) < 1e-6
    o.frames(120)
    assert abs(err()) < 1e-4
    o.art[A["A_JF"]] = 0.05   # a torque on the leader
    worst, qd_worst = 0.0, 0.0
    for f in range(240):
        o.frames(1)
        worst = max(worst, abs(err()))
        qd_worst = max(qd_worst, abs(o.art[A["A_QD"] + 1] + 0.5 * o.art[A["A_QD"]]))
    assert abs(o.art[A["A_Q"]]) > 0.3
    assert worst < 2e-3 and qd_worst < 2e-2


def test_drive_and_spring_settle_at_their_stiffness_weighted_equilibrium():
    """Whole steps (solve + integration): a damped hinge under a drive and a spring comes to rest where their
    torques cancel, q* = (K q_tgt + K_s q_ref) / (K + K_s)."""
    ke, ks, tq, qref = 40.0, 10.0, 0.8, -0.6
    scene = _vertical_hinge(targetKe=ke, targetKd=2.0, targetPos=tq, effortLimit=0.0)
    scene.json["joints"]["spring_stiffness"] = [ks]
    scene.json["joints"]["spring_ref"] = [qref]
    pw = pack_scene(scene)
    o = OracleWorld(pw)
    o.substeps(2400)
    qs = (ke * tq + ks * qref) / (ke + ks)
    assert abs(o.art[pw.A["A_Q"]] - qs) < 1e-4
    assert abs(o.art[pw.A["A_QD"]]) < 1e-3


def test_fp32_oracle_tracks_fp64_for_one_step():
    """Error of one float step against the double reimplementation of the same step (reported figure)."""
    for scene in (workloads.boxes64(), workloads.quadruped_world(1)):
        pw = pack_scene(scene)
        o32, o64 = OracleWorld(pw), OracleWorld(pw, f64=True)
        o32.frames(1)
        o64.frames(1)
        NB = pw.NB
        p32, p64 = o32.poses[:NB * 8].astype(np.float64), o64.poses[:NB * 8]
        assert np.abs(p32 - p64).max() <= 1e-5 * max(1.0, np.abs(p64).max())


def _resting_sphere(warm, relaxation=0.5, r=0.1, z0=0.099):
    b = SceneBuilder(fps=240)
    b.addGroundPlane(gap=0.01)
    s = b.addBody(xform=[0, 0, z0, 0, 0, 0, 1])
    b.addShapeSphere(s, radius=r, gap=0.01)
    scene = b.finalize(solver=dict(substeps=4, iterations=1, relaxation=relaxation))
    return pack_scene(scene, opts=dict(warmStart=warm)), float(np.float32(r))


def test_warm_start_sphere_follows_the_relaxed_single_row_recursion():
    """A sphere resting on the ground, one contact row, one relaxed iteration (omega 0.5): each substep starts the
    row at s times the previous substep's final normal impulse (s = 0, 0.5, 1), applies it in the warm pass, then
    relaxes once: v2 = v + g dt + s l / m; l' = max(s l + omega (-(v2 + rhs) m_eff w - (1 - w) s l), 0);
    v' = v2 + (l' - s l) / m; z' = z + v' dt.  The FP64 oracle must follow this recursion substep by substep,
    and the three scales must give different trajectories (so the warm table is read, scaled and applied)."""
    trajs = {}
    for s in (0.0, 0.5, 1.0):
        pw, r = _resting_sphere(s)
        o = OracleWorld(pw, f64=True)
        SF = pw.h("W_SF_BODY")
        statF = pw.statF.reshape(-1, 4).astype(np.float64)
        im = statF[SF][0]
        cap = statF[SF + 6][3]
        cap = cap if cap > 0 else 1e20
        dt, gz, om = pw.hf("W_DT"), pw.hf("W_GZ"), pw.hf("W_F_OMEGA")
        beta, spec, reg, cfm = pw.hf("W_BETA"), pw.hf("W_F_SPEC"), pw.hf("W_REG"), pw.hf("W_F_CFM")
        assert om == 0.5 and (pw.h("W_WH") > 0) == (s > 0)
        eff = 1.0 / (im + cfm)
        z, v, lam = float(o.poses[pw.h("W_P_BODY") * 4 + 2]), 0.0, 0.0
        zs = []
        for k in range(240):
            o.substeps(1, debug=True)
            assert o.dbg_bufs["ncontacts"][0] == 1
            phi = z - r
            rhs = max(beta * phi / dt, -cap) if phi < 0 else spec * phi / dt
            w = 1.0 if (phi > 0 or reg <= 0) else 1.0 / (1.0 + reg)
            ln = s * lam
            v2 = v + gz * dt + ln * im
            nl = max(ln + om * (-(v2 + rhs) * eff * w - (1.0 - w) * ln), 0.0)
            v = v2 + (nl - ln) * im
            lam = nl
            z = z + v * dt
            P = pw.h("W_P_BODY") * 4
            assert abs(o.poses[P + 2] - z) <= 1e-12, (s, k)
            assert abs(o.vel[2] - v) <= 1e-10, (s, k)
            zs.append(z)
        trajs[s] = np.array(zs)
    assert np.abs(trajs[0.5] - trajs[0.0]).max() > 1e-6 and np.abs(trajs[1.0] - trajs[0.5]).max() > 1e-6


def _incline(theta, warm=1, mu=1.0, anchors=True):
    """A box resting on a static slab tilted by theta about x (static friction holds it)."""
    b = SceneBuilder(fps=240)
    c, s_ = math.cos(theta / 2), math.sin(theta / 2)
    q = [s_, 0, 0, c]
    b.addShapeBox(-1, hx=2, hy=2, hz=0.1, gap=0.01, mu=mu, xform=[0, 0, 0, *q])
    n = np.array([0, -math.sin(theta), math.cos(theta)])
    p = n * (0.1 + 0.1 - 0.0005)
    box = b.addBody(xform=[*p, *q])
    b.addShapeBox(box, hx=0.1, hy=0.1, hz=0.1, gap=0.01, mu=mu)
    scene = b.finalize(solver=dict(substeps=4, iterations=8, **({} if anchors else dict(friction_anchor_beta=None))))
    return pack_scene(scene, opts=dict(warmStart=warm)), n


def test_warm_table_holds_the_equilibrium_impulses_of_a_box_on_an_incline():
    """Box held by static friction on a 20 degree slope: once at rest, the warm table's entries (one per kept
    contact row) sum, as impulses n * normal + tangent vector on the box, to the weight impulse m g dt straight
    up; the normal parts to m g cos(theta) dt and the tangent vectors to m g sin(theta) dt up the slope
    (a swapped or mis-signed tangent basis fails this)."""
    theta = math.radians(20)
    pw, n = _incline(theta)
    o = OracleWorld(pw)
    o.substeps(4 * 240)
    SF = pw.h("W_SF_BODY")
    im = float(pw.statF.reshape(-1, 4)[SF][0])
    m, dt, g = 1.0 / im, pw.hf("W_DT"), -pw.hf("W_GZ")
    assert np.abs(o.vel[:8]).max() < 1e-3
    tab = o.warm_map()
    assert 1 <= len(tab) <= 8
    vals = np.array(list(tab.values()), np.float64)
    # the slab is static geometry (a = -1) and the normal points from it to the box
    J = vals[:, 0].sum() * n + vals[:, 1:].sum(0)
    W = m * g * dt
    assert np.allclose(J, [0, 0, W], atol=0.02 * W), (J, W)
    assert abs(vals[:, 0].sum() - W * math.cos(theta)) < 0.02 * W
    up = np.array([0, math.cos(theta), math.sin(theta)])
    assert abs(vals[:, 1:].sum(0) @ up - W * math.sin(theta)) < 0.02 * W
    assert np.abs(vals[:, 1:] @ n).max() < 1e-6 * W + 1e-9


def _rows_after_prepare(o):
    pw = o.pw
    ROWS = pw.h("W_ROWS")
    nr = int(o.dbg_bufs["nrows"][0])
    return o.dbg_bufs["rows"][:nr * ROWS * 4].reshape(nr, ROWS, 4).astype(np.float64)


def test_warm_start_rows_begin_at_the_scaled_table_values_within_the_cone():
    """After prepare, every contact row starts at the table entry of its key from the previous substep: normal
    s * l and tangent = the scaled stored vector projected on (t1, t2), clamped to mu * normal (checked without
    friction anchors).  With anchors, non-anchor members start at no friction and anchor rows at the carried
    anchor impulse (the anchor table's, rotated by the static side's identity) projected and scaled, clamped to
    mu / na of the members' load.  Incline box, s = 0.5."""
    theta = math.radians(20)
    for anchors in (False, True):
        pw, n = _incline(theta, warm=0.5, anchors=anchors)
        assert pw.h("W_PATCH") == int(anchors)
        o = OracleWorld(pw)
        o.substeps(4 * 60)
        s = pw.hf("W_WARM_SCALE")
        assert s == 0.5
        tab = o.warm_map()
        AH, par = pw.h("W_AH"), int(o.count[0]) % 2
        at = o.anchors_u32()
        akeys, avals = at[par * AH:(par + 1) * AH], at[2 * AH:].reshape(2 * AH, 14)[par * AH:(par + 1) * AH]
        o.substeps(1, debug=True)
        R = _rows_after_prepare(o)
        nr, PO = len(R), pw.h("W_PO")
        assert nr >= 2
        load = sum(R[i, 6, 0] for i in range(nr) if R[i, PO, 3] != 1.0) if anchors else 0.0
        carried = 0
        for i in range(nr):
            key = int(R[i, 3, 3]) | (int(R[i, 4, 3]) << 16)
            v = tab[key].astype(np.float64)
            t1, t2, mu = R[i, 1, :3], R[i, 2, :3], R[i, 2, 3]
            assert abs(R[i, 6, 0] - s * v[0]) <= 1e-6 * abs(v[0])
            if not anchors:
                lt = s * np.array([v[1:] @ t1, v[1:] @ t2])
                lim = mu * s * v[0]
                if np.hypot(*lt) > lim:
                    lt *= lim / np.hypot(*lt)
                assert np.allclose(R[i, 6, 1:3], lt, rtol=1e-5, atol=1e-7 * abs(v[0])), (i, R[i, 6], lt)
                continue
            role = R[i, PO, 3]
            if role == 0.0:
                assert R[i, 6, 1] == 0 and R[i, 6, 2] == 0
            elif role == 2.0:
                pair = int(np.float32(R[i, PO + 3, 3]).view(np.uint32))
                k = int(R[i, PO + 1, 3])
                hit = [j for j in range(AH) if akeys[j] and avals[j, 0] == pair and avals[j, 1] == k]
                if not hit:
                    continue
                il = avals[hit[0], 8:11].view(np.float32).astype(np.float64)
                lw = s * np.array([il @ t1, il @ t2])
                rad = max(mu * load, 0.0)
                if np.hypot(*lw) > rad:
                    lw *= rad / np.hypot(*lw)
                assert np.allclose(R[i, 6, 1:3], lw, rtol=1e-5, atol=1e-7 * load), (i, R[i, 6], lw)
                carried += 1
        assert carried >= 1 or not anchors


def test_nesterov_momentum_sphere_follows_the_accelerated_single_row_recursion():
    """Nesterov momentum (one frictionless-in-effect contact row, omega 0.5, 6 iterations): before sweep it > 0 the
    restart test sums (lam - lam_last - e) * e (below zero: k = 0, else k + 1), beta = (k - 1) / (k + 2) for
    k > 1, the row jumps to y = lam + beta (lam - lam_last) and the change reaches the velocity through the row's
    response, then the sweep relaxes from y.  The FP64 oracle follows this recursion substep by substep, and the
    trajectory differs from the plain one."""
    trajs = {}
    for nest in (False, True):
        b = SceneBuilder(fps=240)
        b.addGroundPlane(gap=0.01)
        s = b.addBody(xform=[0, 0, 0.0985, 0, 0, 0, 1])
        b.addShapeSphere(s, radius=0.1, gap=0.01)
        pw = pack_scene(b.finalize(solver=dict(substeps=4, iterations=6, relaxation=0.5, nesterov_momentum=nest)))
        assert pw.h("W_NEST") == int(nest)
        r = float(np.float32(0.1))
        o = OracleWorld(pw, f64=True)
        SF = pw.h("W_SF_BODY")
        statF = pw.statF.reshape(-1, 4).astype(np.float64)
        im = statF[SF][0]
        dt, gz, om = pw.hf("W_DT"), pw.hf("W_GZ"), pw.hf("W_F_OMEGA")
        beta_b, spec, reg, cfm = pw.hf("W_BETA"), pw.hf("W_F_SPEC"), pw.hf("W_REG"), pw.hf("W_F_CFM")
        eff = 1.0 / (im + cfm)
        P = pw.h("W_P_BODY") * 4
        z, v = float(o.poses[P + 2]), 0.0
        zs, restarts = [], 0
        for k in range(120):
            o.substeps(1, debug=True)
            assert o.dbg_bufs["ncontacts"][0] == 1
            phi = z - r
            rhs = beta_b * phi / dt if phi < 0 else spec * phi / dt
            w = 1.0 if (phi > 0 or reg <= 0) else 1.0 / (1.0 + reg)
            v = v + gz * dt
            lam, last, e, kk = 0.0, 0.0, 0.0, 0.0
            for it in range(6):
                if nest:
                    if it > 0:
                        total = (lam - last - e) * e
                        kk = 0.0 if total < 0 else kk + 1.0
                        restarts += total < 0
                        bt = (kk - 1.0) / (kk + 2.0) if kk > 1 else 0.0
                        y = lam + bt * (lam - last)
                        e = y - lam
                        last = lam
                        if e != 0:
                            lam = y
                            v += e * im
                    else:
                        last, e = lam, 0.0
                nl = max(lam + om * (-(v + rhs) * eff * w - (1.0 - w) * lam), 0.0)
                v += (nl - lam) * im
                lam = nl
            z = z + v * dt
            assert abs(o.poses[P + 2] - z) <= 1e-12, (nest, k)
            assert abs(o.vel[2] - v) <= 1e-10, (nest, k)
            zs.append(z)
        trajs[nest] = np.array(zs)
        if nest:
            assert restarts > 0  # the restart branch is exercised
    assert np.abs(trajs[True] - trajs[False]).max() > 1e-6


def _spinning_ball(tr, spin, R=0.1, mu=0.8):
    """Floating-base ball resting slightly sunk on the ground, spinning about z; its shape has contact torsion
    radius tr.  A small inner ball on a revolute joint about z makes it an articulation without adding to the
    base's spin inertia (the joint decouples the child's z rotation)."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane(gap=0.01, mu=mu)
    s = b.addBody(xform=[0, 0, R - 0.0003, 0, 0, 0, 1])
    b.addShapeSphere(s, radius=R, gap=0.01, mu=mu)
    b.addJointFree(s)
    c = b.addBody(xform=[0, 0, R - 0.0003, 0, 0, 0, 1])
    b.addShapeSphere(c, radius=0.01, gap=0.001)
    b.addJointRevolute(s, c, axis=(0, 0, 1))
    scene = b.finalize(solver=dict(substeps=4, iterations=8, friction_anchor_beta=None))
    scene.json["shapes"][1]["torsion_radius"] = tr
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=True)
    o.art[pw.A["A_QD"] + 5] = spin
    return pw, o


def test_contact_torsion_spins_down_at_radius_times_coulomb_budget():
    """An articulated ball spinning on its contact point: the point does not slide, so sliding friction is zero and
    the spin row alone brakes the spin, at the bound r * mu * N per substep while the ball sinks into the ground
    (a torsion member needs phi <= 0): I_zz dw_z = -r mu N, with N the substep's normal impulse read as
    m (dv_z + g dt), and dw_z measured before the root's angular damping (w' = (w + dw) (1 - c dt)).  With the ball
    above the ground (phi > 0) the spin only damps.  Without a torsion radius the spin only damps."""
    r, mu, w0 = 0.02, 0.8, 6.0
    R = float(np.float32(0.1))
    for tr in (r, 0.0):
        pw, o = _spinning_ball(tr, w0)
        assert pw.h("W_TORS") == int(tr > 0) and pw.NA == 1 and pw.NL == 2
        SF, A = pw.h("W_SF_BODY"), pw.A
        F = pw.statF.reshape(-1, 4).astype(np.float64)
        m, Izz, damp = 1.0 / F[SF][0] + 1.0 / F[SF + 8][0], F[SF + 4][2], F[SF + 3][3]
        dt, gz = pw.hf("W_DT"), pw.hf("W_GZ")
        QD, Q = A["A_QD"], A["A_Q"]
        braked = free = 0
        for k in range(4 * 120):
            v0, w_0, phi = o.art[QD + 2], o.art[QD + 5], o.art[Q + 2] - R
            o.substeps(1)
            v1, w_1 = o.art[QD + 2], o.art[QD + 5]
            N = m * (v1 - v0 - gz * dt)
            dw = w_1 / (1.0 - damp * dt) - w_0
            if tr == 0 or N == 0:  # no torsion radius, or no load this substep: the spin only damps
                assert abs(dw) <= 1e-12 * w0, k
                free += 1
            elif w_1 > 1e-9:
                assert phi <= 0 and N > 0
                # 1e-7: the body record's inverse mass and the link's mass are separately rounded to float32
                assert abs(Izz * dw + r * mu * N) <= 1e-7 * r * mu * N, k
                braked += 1
            else:
                assert abs(w_1) <= 1e-12, k  # stopped within the bound, and held
                free += 1
        assert free > 20 and (braked > 100 or tr == 0)


def _qax(ax, a):
    ax = np.asarray(ax, float) / np.linalg.norm(ax)
    return [*(ax * math.sin(a / 2)), math.cos(a / 2)]


def _qmul(p, q):
    px, py, pz, pw_ = p
    qx, qy, qz, qw = q
    return [pw_ * qx + px * qw + py * qz - pz * qy, pw_ * qy - px * qz + py * qw + pz * qx,
            pw_ * qz + px * qy - py * qx + pz * qw, pw_ * qw - px * qx - py * qy - pz * qz]


def _jointed_box(joints, frame=(0, 0, 0, 0, 0, 0, 1)):
    """A box (COM 0.3 below its frame) hanging from the world origin through `joints`: a list of (type, axes, linear,
    q), chained through tiny-mass intermediate bodies; returns the packed world and the box's body index."""
    from steel_webgpu_amd.scene import JT
    b = SceneBuilder(fps=240)
    parent = -1
    for k, (t, axes, lin, q) in enumerate(joints):
        last = k == len(joints) - 1
        body = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1], **({} if last else dict(mass=1e-6, com=[0, 0, 0], inertia=[1e-9, 0, 0, 0, 1e-9, 0, 0, 0, 1e-9])))
        if last:
            b.addShapeBox(body, hx=0.05, hy=0.08, hz=0.12, xform=[0, 0, -0.3, 0, 0, 0, 1])
        b.addJoint(JT[t], parent, body, parentXform=list(frame) if k == 0 else [0, 0, 0, 0, 0, 0, 1], axes=axes, linear=lin, q=q)
        parent = body
    return pack_scene(b.finalize(solver=dict(substeps=4))), parent


def _track(pw, body, frames, qd=None):
    o = OracleWorld(pw, f64=True)
    if qd is not None:
        o.art[pw.A["A_QD"]:pw.A["A_QD"] + len(qd)] = qd
    P = pw.h("W_P_BODY") * 4
    out = []
    for f in range(frames):
        o.frames(1)
        out.append(np.concatenate([o.poses[P + body * 8:P + body * 8 + 3], o.poses[P + body * 8 + 4:P + body * 8 + 8] * np.sign(o.poses[P + body * 8 + 7])]))
    return np.array(out)


def test_d6_joints_move_like_the_equivalent_ball_universal_and_planar_chains():
    """A D6 joint is Newton's stack of translations then rotations about carried axes.  Three rotations swing a
    hanging box exactly like a ball joint (same initial orientation, intrinsic x then y); two rotations like a
    revolute-revolute chain; two translations and a rotation like a prismatic-prismatic-revolute chain (tiny-mass
    intermediate links; started sliding and turning, qd = (1, 0, 3), the same coordinates in both).  FK, the motion
    subspace and its velocity-product terms all enter: the trajectories agree to the integrators' O(dt) difference
    while the box moves by tenths of a meter."""
    X, Y, Z = [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]
    a0, a1 = 0.5, 0.35
    cases = [
        ([("D6", [X, Y, Z], 0, [a0, a1, 0.0])], [("BALL", [X, Y, Z], 0, _qmul(_qax(X, a0), _qax(Y, a1)))], (0, 0, 0, 0, 0, 0, 1)),
        ([("D6", [X, Y], 0, [a0, a1])], [("REVOLUTE", [X], 0, [a0]), ("REVOLUTE", [Y], 0, [a1])], (0, 0, 0, 0, 0, 0, 1)),
        ([("D6", [X, Y, Z], 2, [0.0, 0.0, a0])], [("PRISMATIC", [X], 1, [0.0]), ("PRISMATIC", [Y], 1, [0.0]), ("REVOLUTE", [Z], 0, [a0])],
         (0, 0, 0, *_qax(X, math.pi / 2))),
    ]
    for i, (d6, chain, frame) in enumerate(cases):
        qd = [1.0, 0.0, 3.0] if i == 2 else None
        pa, ba = _jointed_box(d6, frame)
        pb, bb = _jointed_box(chain, frame)
        ta, tb = _track(pa, ba, 120, qd), _track(pb, bb, 120, qd)
        swing = np.abs(ta[:, :3] - ta[0, :3]).max()
        assert swing > 0.05, (d6, swing)
        err = np.abs(ta - tb).max()
        assert err < 0.01 * swing, (d6, err, swing)


def test_velocity_iterations_sphere_follows_the_two_phase_recursion():
    """velocity_iterations = 2 on a resting sphere (one row, omega 0.5, one position iteration): the position
    sweep's velocity v_s moves the sphere (z' = z + v_s dt); then, from the accumulated impulse, two bias-free
    sweeps (row weight 1, rhs = phi / dt only for a speculative gap still open at v_s, else 0) set the output
    velocity.  The FP64 oracle follows this substep by substep, and differs from the run without them."""
    trajs = {}
    for vi in (0, 2):
        b = SceneBuilder(fps=240)
        b.addGroundPlane(gap=0.01)
        s = b.addBody(xform=[0, 0, 0.0985, 0, 0, 0, 1])
        b.addShapeSphere(s, radius=0.1, gap=0.01)
        pw = pack_scene(b.finalize(solver=dict(substeps=4, iterations=1, relaxation=0.5, velocity_iterations=vi)))
        assert pw.h("W_VELITERS") == vi
        r = float(np.float32(0.1))
        o = OracleWorld(pw, f64=True)
        SF = pw.h("W_SF_BODY")
        im = pw.statF.reshape(-1, 4).astype(np.float64)[SF][0]
        dt, gz, om = pw.hf("W_DT"), pw.hf("W_GZ"), pw.hf("W_F_OMEGA")
        beta_b, spec, reg, cfm = pw.hf("W_BETA"), pw.hf("W_F_SPEC"), pw.hf("W_REG"), pw.hf("W_F_CFM")
        eff = 1.0 / (im + cfm)
        P = pw.h("W_P_BODY") * 4
        z, v = float(o.poses[P + 2]), 0.0
        zs = []
        for k in range(160):
            o.substeps(1, debug=True)
            assert o.dbg_bufs["ncontacts"][0] == 1
            phi = z - r
            rhs = beta_b * phi / dt if phi < 0 else spec * phi / dt
            w = 1.0 if (phi > 0 or reg <= 0) else 1.0 / (1.0 + reg)
            v = v + gz * dt
            lam = max(om * (-(v + rhs) * eff * w), 0.0)
            v += lam * im
            vs = v
            if vi:
                rv = phi / dt if (phi > 0 and phi + dt * v > 1e-6) else 0.0
                for _ in range(vi):
                    nl = max(lam + om * (-(v + rv) * eff), 0.0)
                    v += (nl - lam) * im
                    lam = nl
            z = z + vs * dt
            assert abs(o.poses[P + 2] - z) <= 1e-12, (vi, k)
            assert abs(o.vel[2] - v) <= 1e-10, (vi, k)
            zs.append(v)
        trajs[vi] = np.array(zs)
    assert np.abs(trajs[2] - trajs[0]).max() > 1e-4


def kinematic_turntable(n=3):
    """A kinematic slab spinning at 1.2 rad/s about z (prescribed twist, infinite mass, no gravity) carrying boxes."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane()
    t = b.addBody(xform=[0, 0, 0.05, 0, 0, 0, 1], kinematic=True, twist=(0, 0, 0, 0, 0, 1.2))
    b.addShapeBox(t, hx=0.6, hy=0.6, hz=0.05)
    for k in range(n):
        bx = b.addBody(xform=[0.15 + 0.15 * k, 0.02 * k, 0.15 + 0.001 * k, 0, 0, 0, 1])
        b.addShapeBox(bx, hx=0.05, hy=0.05, hz=0.05)
    return b.finalize(solver=dict(substeps=4))


def test_kinematic_body_keeps_its_prescribed_twist_and_carries_what_rests_on_it():
    """A kinematic body has infinite mass and no gravity: contacts and the boxes' weight leave its twist exactly as
    prescribed, its COM stays put (zero linear velocity) and its yaw advances by w t (quaternion integration of a
    constant rate, to O(dt^2) per step); boxes resting on it are carried round by friction until they turn with
    it, at w r about the axis."""
    pw = pack_scene(kinematic_turntable())
    o = OracleWorld(pw, f64=True)
    dt = pw.hf("W_DT")
    P = pw.h("W_P_BODY") * 4
    c0 = o.poses[P:P + 3].copy()
    n = 4 * 240
    o.substeps(n)
    w = float(np.float32(1.2))
    assert np.array_equal(o.vel[:3], [0, 0, 0]) and o.vel[6] == w and o.vel[4] == 0 and o.vel[5] == 0
    assert np.abs(o.poses[P:P + 3] - c0).max() < 1e-12
    q = o.poses[P + 4:P + 8]
    yaw = 2 * math.atan2(q[2], q[3])
    assert abs(yaw - w * n * dt) < 1e-3
    for k in range(3):
        v = o.vel[(1 + k) * 8:(1 + k) * 8 + 3]
        c = o.poses[P + (1 + k) * 8:P + (1 + k) * 8 + 3]
        expect = w * np.array([-c[1], c[0], 0.0])
        assert np.abs(v - expect).max() < 0.02 * np.linalg.norm(expect) + 1e-3, (k, v, expect)


def test_compliant_contact_settles_at_the_static_spring_deflection():
    """contact_compliance: the normal row is an implicit spring-damper (ke of the two shapes in series, kd summed)
    whose fixed point with the sphere at rest is the static deflection phi = -m g / ke; without compliance the
    rigid contact settles at phi = 0 (Baumgarte)."""
    for compl in (True, False):
        b = SceneBuilder(fps=240)
        b.addGroundPlane(gap=0.01, ke=2e4, kd=50.0)
        s = b.addBody(xform=[0, 0, 0.05, 0, 0, 0, 1])
        b.addShapeSphere(s, radius=0.05, gap=0.01, ke=2e4, kd=50.0)
        pw = pack_scene(b.finalize(solver=dict(substeps=4, iterations=4, contact_compliance=compl)))
        assert pw.h("W_COMPL") == int(compl) and pw.h("W_PATCH") == int(not compl)
        o = OracleWorld(pw, f64=True)
        o.substeps(4 * 480)
        SF = pw.h("W_SF_BODY")
        m = 1.0 / pw.statF.reshape(-1, 4).astype(np.float64)[SF][0]
        g = -pw.hf("W_GZ")
        k = float(np.float32(2e4)) * float(np.float32(2e4)) / (2 * float(np.float32(2e4)))
        phi = o.poses[pw.h("W_P_BODY") * 4 + 2] - float(np.float32(0.05))
        assert abs(o.vel[2]) < 1e-9
        if compl:
            assert abs(phi + m * g / k) <= 1e-6 * m * g / k, (phi, -m * g / k)
        else:
            assert abs(phi) < 1e-6 * m * g / k


def test_drive_rows_reproduce_the_implicit_drive_and_clamp_at_effort_times_dt():
    """drive_mode rows: each substep the drive row starts at zero and, with x = 1 / (1 + c_a / I), c_a = dt (dt K + C_d),
    settles at x (dt C_d qd_tgt + dt K (q_tgt - q)) - x c_a qd (the first visit is already its fixed point), i.e.
    qd' = (I qd + dt (K (q_tgt - q) + C_d qd_tgt)) / (I + dt C_d + dt^2 K): the implicit drive's backward-Euler
    update.  So an isolated hinge follows the same trajectory under both drive modes; with a small effort limit
    the row's impulse saturates at effort * dt, so |qd' - qd| = effort dt / I."""
    ke, kd, tq, tv = 150.0, 4.0, 0.9, -0.5
    runs = {}
    for mode in ("implicit", "rows"):
        scene = _vertical_hinge(targetKe=ke, targetKd=kd, targetPos=tq, targetVel=tv, effortLimit=0.0)
        scene.json["solver"]["drive_mode"] = mode
        pw = pack_scene(scene)
        assert pw.h("W_DRIVE_ROWS") == int(mode == "rows")
        o = OracleWorld(pw, f64=True)
        qs = []
        for k in range(240):
            o.substeps(1)
            qs.append((o.art[pw.A["A_Q"]], o.art[pw.A["A_QD"]]))
        runs[mode] = np.array(qs)
    assert np.abs(runs["rows"] - runs["implicit"]).max() < 1e-9
    assert abs(runs["rows"][-1, 0] - runs["rows"][0, 0]) > 0.1  # it moved
    # saturated: effort 0.05 N m
    eff = 0.05
    scene = _vertical_hinge(targetKe=ke, targetKd=kd, targetPos=tq, targetVel=tv, effortLimit=eff)
    scene.json["solver"]["drive_mode"] = "rows"
    H, _ = _crba(scene)
    pw = pack_scene(scene)
    o = OracleWorld(pw, f64=True)
    dt = pw.hf("W_DT")
    I = H[0, 0]
    for k in range(20):
        qd0 = o.art[pw.A["A_QD"]]
        o.substeps(1)
        assert abs(abs(o.art[pw.A["A_QD"]] - qd0) - _f32(eff) * dt / I) <= 1e-6 * _f32(eff) * dt / I, k


def _three_bar(loop_type):
    """A bar hinged (about y) at the world pivot (0, 0, 1) and a second bar on a ball joint at its end, closed to the
    world point (0.5, 0, 0.5) by a loop joint at the second bar's foot: a triangle, rigid under gravity only
    through the closure."""
    from steel_webgpu_amd.scene import JT
    b = SceneBuilder(fps=240)
    A = b.addBody(xform=[0.25, 0, 1, 0, 0, 0, 1])
    b.addShapeBox(A, hx=0.25, hy=0.02, hz=0.02)
    B = b.addBody(xform=[0.5, 0, 0.75, 0, 0, 0, 1])
    b.addShapeBox(B, hx=0.02, hy=0.02, hz=0.25)
    b.addJointRevolute(-1, A, axis=[0, 1, 0], parentXform=[0, 0, 1, 0, 0, 0, 1], childXform=[-0.25, 0, 0, 0, 0, 0, 1])
    b.addJointBall(A, B, parentXform=[0.25, 0, 0, 0, 0, 0, 1], childXform=[0, 0, 0.25, 0, 0, 0, 1])
    if loop_type is not None:
        axes = [[0, 1, 0]] if loop_type == "REVOLUTE" else [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
        b.addJoint(JT[loop_type], -1, B, parentXform=[0.5, 0, 0.5, 0, 0, 0, 1], childXform=[0, 0, -0.25, 0, 0, 0, 1], axes=axes,
                   q=[0.0] if loop_type == "REVOLUTE" else [0, 0, 0, 1])
    return pack_scene(b.finalize(solver=dict(substeps=4, iterations=12))), B


def test_loop_closures_hold_the_closed_chain_and_its_hinge_axis():
    """A ball or revolute loop joint closes a two-bar chain to the world (Baumgarte-stabilized bilateral rows at
    the anchors' midpoint): under gravity the foot stays on its world point (gap under 1 mm after 2 s, from exactly
    0).  The second bar starts spinning about its own axis (2 rad/s through its ball joint): a revolute closure's
    two hinge rows stop that and keep the bar's y axis on the world's, a ball closure lets it turn.  Without the
    closure the chain collapses."""
    for lt in ("BALL", "REVOLUTE", None):
        pw, B = _three_bar(lt)
        assert pw.h("W_NLOOP") == (0 if lt is None else 1)
        o = OracleWorld(pw, f64=True)
        o.art[pw.A["A_QD"] + 3] = 2.0  # the ball joint's rate about z
        o.substeps(4 * 480)
        P = pw.h("W_P_BODY") * 4
        c = o.poses[P + B * 8:P + B * 8 + 3]
        q = o.poses[P + B * 8 + 4:P + B * 8 + 8]
        foot = c + _rot(q) @ np.array([0, 0, -0.25])
        gap = np.linalg.norm(foot - [0.5, 0, 0.5])
        tilt = np.linalg.norm(np.cross(_rot(q) @ [0, 1, 0], [0, 1, 0]))
        if lt is None:
            assert gap > 0.1
        else:
            assert gap < 1e-3, (lt, gap)
            assert (tilt < 1e-3) if lt == "REVOLUTE" else (tilt > 0.1), (lt, tilt)


def _spinning_free_ball(condim, spin=6.0, roll=0.0, mut=0.02, mur=0.01, R=0.1):
    b = SceneBuilder(fps=240)
    b.addGroundPlane(gap=0.01)
    s = b.addBody(xform=[0, 0, R - 0.0003, 0, 0, 0, 1], twist=(roll * R, 0, 0, 0, roll, spin))
    b.addShapeSphere(s, radius=R, gap=0.01)
    scene = b.finalize(solver=dict(substeps=4, iterations=8))
    for sh in scene.json["shapes"]:
        sh.update(condim=condim, mu_torsional=mut, mu_rolling=mur)
    scene.json["bodies"]["angular_damping"] = [0.0]
    return pack_scene(scene)


def test_free_body_spin_row_brakes_at_mu_t_times_load_and_rolling_rows_stop_a_rolling_ball():
    """condim >= 4 on a free body: a ball spinning on its contact point is braked only by the spin row, within
    mu_t N per substep: I_zz dw_z = -mu_t N (N read from the vertical momentum change), until it stops exactly.
    condim 6 adds rolling rows (mu_r N): a rolling ball comes to rest, while with condim 3 it keeps rolling."""
    pw = _spinning_free_ball(4)
    assert pw.h("W_ANG") == 1 and pw.h("W_PO") == 34
    o = OracleWorld(pw, f64=True)
    SF = pw.h("W_SF_BODY")
    F = pw.statF.reshape(-1, 4).astype(np.float64)
    m, Izz = 1.0 / F[SF][0], F[SF + 4][2]
    dt, gz = pw.hf("W_DT"), pw.hf("W_GZ")
    mut = float(np.float32(0.02))
    braked = stopped = 0
    for k in range(4 * 240):
        v0, w0 = o.vel[2], o.vel[6]
        o.substeps(1)
        N = m * (o.vel[2] - v0 - gz * dt)
        if N == 0:
            assert o.vel[6] == w0
        elif o.vel[6] > 1e-9:
            assert abs(Izz * (o.vel[6] - w0) + mut * N) <= 1e-7 * mut * N, k
            braked += 1
        else:
            assert abs(o.vel[6]) <= 1e-12
            stopped += 1
    assert braked > 50 and stopped > 20
    finals = {}
    for condim in (6, 3):
        pw = _spinning_free_ball(condim, spin=0.0, roll=3.0)
        o = OracleWorld(pw, f64=True)
        o.substeps(4 * 480)
        finals[condim] = np.abs(o.vel[:8]).max()
    assert finals[6] < 1e-3 and finals[3] > 0.1, finals


def fader_rig():
    """A slider (prismatic, pushed along +x by gravity on a tilted rail) and a velocity-driven wheel, coupled by
    the reference's fader controller: the wheel's target speed is max_speed * clamp(q_slider / travel, 0, 1)."""
    b = SceneBuilder(fps=240)
    s = b.addBody(xform=[0, 0, 1, 0, 0, 0, 1])
    b.addShapeBox(s, hx=0.03, hy=0.03, hz=0.03)
    b.addJointPrismatic(-1, s, axis=[1, 0, 0], parentXform=[0, 0, 1, 0, 0.2588190, 0, 0.9659258])
    w = b.addBody(xform=[1, 0, 1, 0, 0, 0, 1])
    b.addShapeCylinder(w, radius=0.1, halfHeight=0.02)
    b.addJointRevolute(-1, w, axis=[0, 0, 1], parentXform=[1, 0, 1, 0, 0, 0, 1], targetKd=5.0)
    scene = b.finalize(solver=dict(substeps=4))
    scene.json["control"] = dict(type="fader", fader_q=0, wheel_v=1, travel=0.24, max_speed=6.283185307179586)
    return scene


def test_fader_controller_sets_the_wheel_target_speed_from_the_slider_each_substep():
    """The fader controller (the reference's demo update_control): every substep, before the drives act, the
    wheel's target velocity becomes max_speed * clamp(q_slider / travel, 0, 1) of the slider coordinate at the
    substep's start; the velocity-driven wheel follows it up to max_speed once the slider passes the travel."""
    pw = pack_scene(fader_rig())
    assert pw.h("W_CTRL_KIND") == 1
    o = OracleWorld(pw, f64=True)
    A = pw.A
    qa, vb = A["A_Q"] + pw.h("W_CTRL_A"), A["A_TGTV"] + pw.h("W_CTRL_B")
    travel, vmax = float(np.float32(0.24)), float(np.float32(6.283185307179586))
    seen = set()
    for k in range(4 * 240):
        q0 = o.art[qa]
        o.substeps(1)
        assert o.art[vb] == vmax * min(max(q0 / travel, 0.0), 1.0), k
        seen.add(round(min(max(q0 / travel, 0.0), 1.0), 1))
    assert 0.0 in seen and 1.0 in seen and len(seen) > 4
    assert abs(o.art[A["A_QD"] + pw.h("W_CTRL_B")] - vmax) < 0.05 * vmax


def test_coulomb_newton_block_meets_the_coulomb_law_and_bisection_the_normal_target():
    """friction_mode blocks (wgsl.js:2583-2653) on random well-conditioned 3x3 Delassus blocks G, start impulses
    lam and live velocities u (post-velocity u' = u + G (lam' - lam)):
    - coulomb_newton: lam' is a Coulomb point: either separating (lam' = 0, u'_n >= target) or touching with
      u'_n = target and either sticking (u'_t = 0, |lam'_t| <= mu lam'_n) or sliding (|lam'_t| = mu lam'_n,
      lam'_t opposite u'_t);
    - bisection and bisection_desaxce: lam'_n >= 0, |lam'_t| <= mu lam'_n, and u'_n meets the (De Saxce-shifted
      for mode 2) target when the root lies inside the search bracket [0, max(2 lam_n, (target - u_n) / G_nn +
      lam_n, 1)] (a coupled block can put it beyond; the search then returns the bracket's end)."""
    L = lib()
    L.oracle_contact_block_f64.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(11)
    kinds = {"stick": 0, "slide": 0, "sep": 0}
    for trial in range(400):
        M = rng.normal(size=(3, 3))
        G = M @ M.T + 0.3 * np.eye(3)
        lam = np.array([abs(rng.normal()), 0.1 * rng.normal(), 0.1 * rng.normal()])
        u = rng.normal(size=3)
        mu = rng.uniform(0.1, 1.2)
        vt = rng.uniform(-0.2, 0.2)
        for mode in (1, 2, 3):
            inp = np.array([u[0], u[1], u[2], G[0, 0], G[0, 1], G[0, 2], G[1, 1], G[1, 2], G[2, 2], vt, mu, *lam])
            out = np.zeros(3)
            L.oracle_contact_block_f64(mode, inp.ctypes.data, out.ctypes.data)
            up = u + G @ (out - lam)
            lt = np.hypot(out[1], out[2])
            assert out[0] >= 0 and lt <= mu * out[0] * (1 + 1e-9) + 1e-12, (trial, mode)
            if mode == 3:
                if out[0] == 0:
                    assert np.all(out == 0) and up[0] >= vt - 1e-9, trial
                    kinds["sep"] += 1
                    continue
                # the root search stops at |f| < 1e-6 (1 + phi0): residuals of that order remain
                assert abs(up[0] - vt) <= 2e-5 * (1 + abs(vt) + out[0] * G[0, 0]), (trial, up[0], vt)
                vtn = np.hypot(up[1], up[2])
                if lt < mu * out[0] * (1 - 1e-6):
                    assert vtn <= 2e-5 * (1 + out[0] * G[0, 0]), (trial, vtn)
                    kinds["stick"] += 1
                else:
                    assert vtn <= 2e-5 * (1 + out[0] * G[0, 0]) or np.dot(out[1:], up[1:]) <= -0.999 * lt * vtn, trial
                    kinds["slide"] += 1
            else:
                target = vt + (mu * np.hypot(u[1], u[2]) if mode == 2 else 0.0)
                hi = max(2 * lam[0], (target - u[0]) / G[0, 0] + lam[0], 1.0)
                if 0 < out[0] < 0.999 * hi:
                    assert abs(up[0] - target) <= 1e-4 * (1 + abs(target) + out[0] * G[0, 0]), (trial, mode, up[0], target)
                    kinds["bis"] = kinds.get("bis", 0) + 1
    assert min(kinds.values()) > 10, kinds


def test_articulated_cone_block_solves_the_ball_constrained_quadratic():
    """The elliptic spin/rolling block (wgsl.js:2655-2925) with omega 1 and no creep: on random SPD 5x5 responses H,
    residuals g, coefficients mu, start impulses lam and loads N, the new impulses lam' minimize
    Q(lam') = 0.5 (lam' - lam)' H (lam' - lam) + g' (lam' - lam) over |lam' / mu| <= N: either interior with zero
    gradient, or on the ellipsoid with the gradient g + H (lam' - lam) = -nu lam' / mu^2, nu >= 0 (KKT)."""
    L = lib()
    L.oracle_ell_block_f64.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double] * 3 + [ctypes.c_void_p]
    rng = np.random.default_rng(5)
    seen = {"interior": 0, "boundary": 0}
    for trial in range(300):
        M = rng.normal(size=(5, 5))
        H = M @ M.T + 0.05 * np.eye(5)
        g = rng.normal(size=5) * rng.uniform(0.1, 3)
        mu = rng.uniform(0.05, 1.0, size=5)
        lam = rng.normal(size=5) * 0.1
        N = rng.uniform(0.1, 2.0)
        out = np.zeros(5)
        L.oracle_ell_block_f64(H.ctypes.data, g.ctypes.data, mu.ctypes.data, lam.ctypes.data, N, 1.0, 0.0, out.ctypes.data)
        y = out / mu
        ny = np.linalg.norm(y)
        assert ny <= N * (1 + 1e-4), (trial, ny, N)
        grad = g + H @ (out - lam)
        if ny < N * (1 - 1e-4):
            assert np.linalg.norm(grad) <= 1e-6 * (1 + np.linalg.norm(g)), trial
            seen["interior"] += 1
        else:
            nrm = out / mu ** 2
            nu = -grad @ nrm / (nrm @ nrm)
            assert nu >= -1e-8, trial
            assert np.linalg.norm(grad + nu * nrm) <= 2e-3 * (np.linalg.norm(grad) + 1e-9), (trial, np.linalg.norm(grad + nu * nrm))
            seen["boundary"] += 1
    assert min(seen.values()) >= 5, seen
