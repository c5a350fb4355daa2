"""Pins of the oracle's default-path functions that have no closed form of their own, against independent
computations: the articulated-body functions against a composite-rigid-body mass matrix with forward kinematics
written here; grouping and body-pair reduction against brute force over the same sorted contacts; the generic
manifold against invariants and a brute-force separating-axis search; friction patches by a box held on an
incline.  Each check was confirmed to fail on a planted mistake in the oracle before it was kept."""
import math

import numpy as np
import pytest

from oracle_bind import OracleWorld, art_probe
from steel_webgpu_amd import SceneBuilder, pack, workloads
from steel_webgpu_amd.pack import pack_scene


# ---------------------------------------------------------------------------------------------------------------
# Independent kinematics and mass matrix of fixed-base revolute trees
# ---------------------------------------------------------------------------------------------------------------

def _quat_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _axis_angle(a, t):
    """Rodrigues rotation matrix about unit axis a by angle t."""
    a = np.asarray(a, np.float64) / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def _unit_quat(rng):
    q = rng.normal(size=4)
    return list(q / np.linalg.norm(q))


def revolute_tree(rng, n, branching=0.0, q_range=1.0, rotated_frames=True):
    """Fixed-base tree of n capsule links on revolute joints with random unit axes, random joint angles and
    (optionally) rotated parent joint frames; with branching > 0 a link's parent is a random earlier link."""
    b = SceneBuilder(fps=240)
    bodies, lens = [], []
    for i in range(n):
        L = float(rng.uniform(0.1, 0.3))
        bd = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1])
        b.addShapeCapsule(bd, radius=0.03, halfHeight=L / 2, xform=[0, 0, L / 2, 0, 0, 0, 1])
        if i == 0:
            parent, off = -1, 0.0
        else:
            k = int(rng.integers(0, i)) if (branching > 0 and rng.uniform() < branching) else i - 1
            parent, off = bodies[k], lens[k]
        ax = rng.normal(size=3)
        rq = _unit_quat(rng) if rotated_frames else [0, 0, 0, 1]
        b.addJointRevolute(parent, bd, axis=list(ax / np.linalg.norm(ax)), parentXform=[0, 0, off, *rq],
                           q=float(rng.uniform(-q_range, q_range)))
        bodies.append(bd)
        lens.append(L)
    return b.finalize(solver=dict(substeps=1))


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def tree_kinematics(scene):
    """World rotation, joint origin, world axis and COM of every joint's child body (joint order), computed from
    the scene's joint arrays (rounded to float32, as the packed world holds them): child = parent * X_p *
    rot(axis, q) * X_c^-1, with X_c the identity."""
    S, J = scene.json, scene.json["joints"]
    n = J["count"]
    R_body, p_body = {}, {}
    out = []
    for j in range(n):
        assert J["type"][j] == 1
        par, ch = J["parent"][j], J["child"][j]
        Rp, pp = (np.eye(3), np.zeros(3)) if par < 0 else (R_body[par], p_body[par])
        xp = _f32(J["X_p"][j])
        o = pp + Rp @ xp[:3]
        Rj = Rp @ _quat_mat(xp[3:7])
        a = _f32(J["axis"][J["qd_start"][j]])
        a = a / np.linalg.norm(a)
        xc = J["X_c"][j]
        assert np.allclose(xc, [0, 0, 0, 0, 0, 0, 1])
        R = Rj @ _axis_angle(a, float(_f32(J["q"][J["q_start"][j]])))
        R_body[ch], p_body[ch] = R, o
        com = o + R @ _f32(S["bodies"]["com"][ch])
        out.append(dict(parent=par, child=ch, R=R, origin=o, axis=Rj @ a, com=com))
    return out


def _ancestors(kin):
    """For each joint, the set of joints on its path to the root (itself included)."""
    by_child = {k["child"]: j for j, k in enumerate(kin)}
    anc = []
    for j, k in enumerate(kin):
        s, p = [j], k["parent"]
        while p >= 0:
            s.append(by_child[p])
            p = kin[by_child[p]]["parent"]
        anc.append(set(s))
    return anc


def com_jacobian(kin, anc, j):
    """6 x n Jacobian of joint j's child body twist at its COM, rows (linear velocity; angular velocity)."""
    n = len(kin)
    Jm = np.zeros((6, n))
    for k in anc[j]:
        w = kin[k]["axis"]
        Jm[:3, k] = np.cross(w, kin[j]["com"] - kin[k]["origin"])
        Jm[3:, k] = w
    return Jm


def point_jacobian(kin, anc, j, p):
    """3 x n Jacobian of the velocity of world point p carried by joint j's child body."""
    n = len(kin)
    Jm = np.zeros((3, n))
    for k in anc[j]:
        Jm[:, k] = np.cross(kin[k]["axis"], p - kin[k]["origin"])
    return Jm


def crba(scene, kin, anc):
    """Joint-space mass matrix: sum over bodies of m Jv^T Jv + Jw^T I_world Jw (joint order)."""
    S = scene.json
    n = len(kin)
    H = np.zeros((n, n))
    for j, k in enumerate(kin):
        b = k["child"]
        m = float(_f32(S["bodies"]["mass"][b]))
        Iw = k["R"] @ _f32(S["bodies"]["inertia"][b]).reshape(3, 3) @ k["R"].T
        Jm = com_jacobian(kin, anc, j)
        H += m * Jm[:3].T @ Jm[:3] + Jm[3:].T @ Iw @ Jm[3:]
    return H


def _art_order(pw, scene):
    """perm[d] = joint index of articulation dof d; link_of_joint[j] = link index of joint j."""
    J = scene.json["joints"]
    perm = np.zeros(pw.ND, np.int64)
    for j in range(J["count"]):
        perm[pw.dofMap[J["qd_start"][j]]] = j
    link_of_joint = {}
    for l, lk in enumerate(pw.links):
        link_of_joint[lk["joint"]] = l
    return perm, link_of_joint


def _setup(scene):
    pw = pack_scene(scene)
    kin = tree_kinematics(scene)
    anc = _ancestors(kin)
    H = crba(scene, kin, anc)
    perm, lj = _art_order(pw, scene)
    Ha = H[np.ix_(perm, perm)]  # articulation dof order
    return pw, kin, anc, H, Ha, perm, lj


TREES = [("chain32", dict(n=32, branching=0.0)), ("chain7_rotated", dict(n=7, branching=0.0)),
         ("tree12", dict(n=12, branching=0.6)), ("tree20", dict(n=20, branching=0.4))]


@pytest.mark.parametrize("name,kw", TREES)
def test_link_frames_match_independent_forward_kinematics(name, kw):
    scene = revolute_tree(np.random.default_rng(11), **kw)
    pw, kin, anc, H, Ha, perm, lj = _setup(scene)
    fr = art_probe(pw, 4, nout=10 * pw.NL).reshape(pw.NL, 10)
    for j, k in enumerate(kin):
        l = lj[j]
        # the packed joint frames are float32: agreement to float32 resolution
        np.testing.assert_allclose(fr[l, 0:3], k["origin"], atol=1e-6)
        np.testing.assert_allclose(_quat_mat(fr[l, 3:7]), k["R"], atol=1e-6)
        np.testing.assert_allclose(fr[l, 7:10], k["com"], atol=1e-6)


@pytest.mark.parametrize("name,kw", TREES)
def test_link_response_is_j_hinv_jt_and_dof_hinv_is_the_diagonal_of_hinv(name, kw):
    """W2 of every link (the unsplit response to a unit force or torque at its COM, artwgsl.js:431-510) equals
    J_l H^-1 J_l^T with H from CRBA; dof_hinv (artwgsl.js:513-518) equals diag(H^-1); with split counts of one
    the split response W equals W2."""
    scene = revolute_tree(np.random.default_rng(3), **kw)
    pw, kin, anc, H, Ha, perm, lj = _setup(scene)
    Hinv = np.linalg.inv(H)
    scale = np.abs(Hinv).max()
    for j in range(len(kin)):
        l = lj[j]
        Jm = com_jacobian(kin, anc, j)
        want = Jm @ Hinv @ Jm.T
        out = art_probe(pw, 0, l)
        W2, W = out[:36].reshape(6, 6), out[36:].reshape(6, 6)
        wscale = np.abs(want).max()
        np.testing.assert_allclose(W2, want, atol=1e-5 * wscale, err_msg=f"link {l}")
        np.testing.assert_allclose(W, W2, atol=0)
        h = art_probe(pw, 1, l, 0, nout=1)[0]
        assert abs(h - Hinv[j, j]) <= 1e-5 * scale, (l, h, Hinv[j, j])


def test_split_response_adds_the_shared_joints_terms():
    """The split response W (Steel's Jacobi split) is W2 plus, for every path joint shared by k contact links,
    (k - 1) times a positive semidefinite term; so W - W2 is PSD, zero for counts of one, and grows linearly."""
    scene = revolute_tree(np.random.default_rng(5), n=9)
    pw, kin, anc, H, Ha, perm, lj = _setup(scene)
    l = lj[8]
    np_ = 9
    base = art_probe(pw, 0, l, inp=np.ones(np_))
    two = art_probe(pw, 0, l, inp=np.full(np_, 2.0))
    three = art_probe(pw, 0, l, inp=np.full(np_, 3.0))
    d2 = two[36:].reshape(6, 6) - base[:36].reshape(6, 6)
    d3 = three[36:].reshape(6, 6) - base[:36].reshape(6, 6)
    np.testing.assert_allclose(d3, 2 * d2, atol=1e-12 * max(1, np.abs(d2).max()))
    ev = np.linalg.eigvalsh(0.5 * (d2 + d2.T))
    assert ev.min() >= -1e-10 * max(1.0, ev.max()) and ev.max() > 0


@pytest.mark.parametrize("name,kw", TREES)
def test_art_solve_gen_is_hinv_tau_and_gen_jacobian_is_the_point_jacobian(name, kw):
    """art_solve_gen (artwgsl.js:557-596) returns H^-1 tau; add_gen_jac (artwgsl.js:525-540) returns the rows
    d . dp/dq of a point carried by a link (geometric Jacobian w_k x (p - o_k) over the link's root path)."""
    rng = np.random.default_rng(17)
    scene = revolute_tree(np.random.default_rng(7), **kw)
    pw, kin, anc, H, Ha, perm, lj = _setup(scene)
    NDA = pw.h("W_NDA")
    for _ in range(4):
        tau = rng.normal(size=pw.ND)
        inp = np.zeros(NDA)
        inp[:pw.ND] = tau
        y = art_probe(pw, 2, 0, inp=inp, nout=NDA)[:pw.ND]
        want = np.linalg.solve(Ha, tau)
        np.testing.assert_allclose(y, want, atol=1e-5 * np.abs(want).max())
    inv = np.argsort(perm)
    for j in rng.choice(len(kin), size=min(6, len(kin)), replace=False):
        p = kin[j]["com"] + rng.normal(size=3) * 0.1
        D = rng.normal(size=(3, 3))
        out = art_probe(pw, 3, lj[j], inp=np.concatenate([p, D.ravel()]), nout=3 * NDA).reshape(3, NDA)[:, :pw.ND]
        want = D @ point_jacobian(kin, anc, j, p)  # joint order
        np.testing.assert_allclose(out[:, inv], want, atol=1e-6 * max(1.0, np.abs(want).max()))


def test_dense_rows_hold_j_and_hinv_jt_of_the_self_contacts_of_a_32_link_arm():
    """Prepare's same-articulation rows (wgsl.js:2365-2444): for each self-contact of the arm, J's three rows are
    n, t1, t2 dotted with the relative velocity of the contact point carried by the two links (independent
    geometric Jacobian), and Y = H^-1 J^T with CRBA H; the row's normal mass is 1 / (J H^-1 J^T)_00."""
    rng = np.random.default_rng(23)
    scene = revolute_tree(rng, n=32, q_range=1.2, rotated_frames=False)
    pw, kin, anc, H, Ha, perm, lj = _setup(scene)
    o = OracleWorld(pw, f64=True)
    o.substeps(1, debug=True)
    nd = int(o.dbg_bufs["meta"][3])
    assert nd >= 3, nd
    NDA, DPS, DHDR, A_DENSE = pw.h("W_NDA"), pw.h("W_DPS"), pw.h("W_DHDR"), pw.A["A_DENSE"]
    C = o.dbg_bufs["contacts"][:int(o.dbg_bufs["ncontacts"][0]) * 24].reshape(-1, 24)
    key_to_contact = {(int(c[10]), int(c[11])): c for c in C}
    joint_of_body = {k["child"]: j for j, k in enumerate(kin)}
    inv = np.argsort(perm)
    Hinv = np.linalg.inv(Ha)
    checked = 0
    for slot in range(min(nd, pw.h("W_MAXDENSE"))):
        blk = o.art[A_DENSE + slot * DPS:A_DENSE + (slot + 1) * DPS]
        c = key_to_contact[(int(blk[1]), int(blk[2]))]
        a, b = int(c[8]), int(c[9])
        p = c[0:3].astype(np.float64)
        t1, t2, n = blk[12:15], blk[15:18], blk[18:21]
        np.testing.assert_allclose(n, c[4:7], atol=1e-6)
        D = np.stack([n, t1, t2])
        want_J = D @ (point_jacobian(kin, anc, joint_of_body[b], p) - point_jacobian(kin, anc, joint_of_body[a], p))
        Jgot = blk[DHDR:DHDR + 3 * NDA].reshape(3, NDA)[:, :pw.ND]
        Ygot = blk[DHDR + 3 * NDA:DHDR + 6 * NDA].reshape(3, NDA)[:, :pw.ND]
        # the contact point is stored as float32 in the debug record: compare at that resolution
        np.testing.assert_allclose(Jgot[:, inv], want_J, atol=2e-6)
        np.testing.assert_allclose(Ygot, (Hinv @ Jgot.T).T, atol=1e-5 * np.abs(Ygot).max())
        K = Jgot @ Ygot.T
        if K[0, 0] > 1e-6 * np.trace(K):
            assert abs(blk[5] - 1.0 / (K[0, 0] + pw.hf("W_F_CFM"))) <= 1e-9 * abs(blk[5])
        checked += 1
    assert checked >= 3


# ---------------------------------------------------------------------------------------------------------------
# Grouping and body-pair reduction, brute force over the sorted contacts
# ---------------------------------------------------------------------------------------------------------------

def _pair_key(c):
    return (int(c[8]) + 1) | ((int(c[9]) + 1) << 16)


def _reduction_case(scene, substeps):
    pw = pack_scene(scene)
    o = OracleWorld(pw)
    o.substeps(substeps - 1)
    o.substeps(1, debug=True)
    nc = int(o.dbg_bufs["ncontacts"][0])
    C = o.dbg_bufs["contacts"][:nc * 24].reshape(nc, 24)
    red = o.dbg_bufs["red"][:nc * 8].reshape(nc, 8)
    redf = o.dbg_bufs["redf"][:nc * 4].reshape(nc, 4)
    nr = int(o.dbg_bufs["nrows"][0])
    info = o.dbg_bufs["rowinfo"][:nr * 4].reshape(nr, 4)
    return pw, C, red, redf, info


def compound_tiles(n=6, grid=3, seed=4):
    """Bodies made of grid x grid small boxes (one rigid tile each), dropped tilted onto the ground and onto each
    other: many contacts per body pair, of which the reduction keeps the extremes and the deepest."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder(fps=240)
    b.addGroundPlane(mu=0.6)
    for i in range(n):
        q = _unit_quat(rng) if i % 2 else [0, 0, 0, 1]
        bd = b.addBody(xform=[float(rng.uniform(-0.2, 0.2)), float(rng.uniform(-0.2, 0.2)), 0.1 + 0.15 * i, *q])
        for gx in range(grid):
            for gy in range(grid):
                b.addShapeBox(bd, hx=0.02, hy=0.02, hz=0.02, mu=0.6, xform=[0.04 * (gx - 1), 0.04 * (gy - 1), 0, 0, 0, 0, 1])
    return b.finalize(solver=dict(substeps=4, iterations=12))


def tilted_tiles(grid=4, angle=0.05):
    """One body of grid x grid small boxes tilted by `angle` about x, its lowest corners 1 mm into the ground: the
    low side touches (gap <= 0), the high side is speculative (gap > 0), in one reduction group."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane(mu=0.6)
    c, s_ = math.cos(angle / 2), math.sin(angle / 2)
    half = 0.02 * grid
    z = 0.02 * math.cos(angle) + half * math.sin(angle) - 0.001
    bd = b.addBody(xform=[0, 0, z, s_, 0, 0, c])
    for gx in range(grid):
        for gy in range(grid):
            b.addShapeBox(bd, hx=0.02, hy=0.02, hz=0.02, mu=0.6,
                          xform=[0.04 * (gx - (grid - 1) / 2), 0.04 * (gy - (grid - 1) / 2), 0, 0, 0, 0, 1])
    return b.finalize(solver=dict(substeps=4, iterations=12))


REDUCTION_SCENES = [("tiles", lambda: compound_tiles(), 240), ("tilted_tiles", lambda: tilted_tiles(), 1), ("boxes64", lambda: workloads.boxes64(), 240),
                    ("quadruped", lambda: workloads.quadruped_world(5), 200), ("hulls", lambda: workloads.arm_hulls(nlinks=4, nhulls=40), 300)]


@pytest.mark.parametrize("name,make,substeps", REDUCTION_SCENES)
def test_units_are_the_runs_of_equal_pair_keys(name, make, substeps):
    """Grouping (wgsl.js:1889-2028): over the sorted contacts, skipping same-articulation ones, every maximal run
    of equal body-pair key is one unit, units are numbered in run order, and a unit lists its run's contacts."""
    pw, C, red, redf, info = _reduction_case(make(), substeps)
    body_art = {}
    SFB = pw.statF.reshape(-1, 4)[pw.h("W_SF_BODY"):]
    for b in range(pw.NB):
        kind, art = int(SFB[b * 8][1]), int(SFB[b * 8][2])
        body_art[b] = art if kind == 2 else None
    runs = []
    for c in range(len(C)):
        a, b = int(C[c][8]), int(C[c][9])
        if a >= 0 and b >= 0 and body_art[a] is not None and body_art[a] == body_art[b]:
            continue  # same-articulation contacts become dense rows, not units
        if runs and _pair_key(C[runs[-1][-1]]) == _pair_key(C[c]):
            runs[-1].append(c)
        else:
            runs.append([c])
    units = {}
    for c in range(len(C)):
        u = int(red[c, 0])
        if u != 0xFFFFFFFF:
            units.setdefault(u, []).append(c)
    assert len(units) == len(runs) >= (1 if name == "tilted_tiles" else 4)
    for u, r in enumerate(runs):
        assert sorted(units[u]) == r
        assert all(_pair_key(C[c]) == _pair_key(C[r[0]]) for c in r)
    # rows carry their unit's pair key
    for r in range(len(info)):
        u = int(info[r, 2])
        assert int(info[r, 3]) == _pair_key(C[runs[u][0]])


@pytest.mark.parametrize("name,make,substeps", REDUCTION_SCENES)
def test_reduction_keeps_exactly_the_brute_force_slot_winners(name, make, substeps):
    """Body-pair reduction (wgsl.js:1459-1503, 2029-2105): within a unit, contacts of equal (group pair, normal bin
    and cell) form a group; a group's 13 slots are won by the largest score (6 extremes along 60 degree
    directions, the deepest, 6 touching extremes for gap <= 0), ties to the larger position identity; the kept
    contacts are exactly the winners of at least one slot, in listing order.  Checked by brute force over the
    exported per-contact inputs, and the inputs against an FP64 recomputation of bin, frame and position."""
    pw, C, red, redf, info = _reduction_case(make(), substeps)
    listed = [c for c in range(len(C)) if red[c, 0] != 0xFFFFFFFF]
    groups = {}
    for c in listed:
        groups.setdefault((int(red[c, 0]), int(red[c, 2]), int(red[c, 3]), int(red[c, 4])), []).append(c)
    from oracle_bind import lib as olib
    ang = [np.float32(k) * np.float32(1.0471975511965976) for k in range(6)]
    cs = [np.float32(olib().oracle_cosf(float(a))) for a in ang]
    sn = [np.float32(olib().oracle_sinf(float(a))) for a in ang]

    def fflip(x):
        u = int(np.float32(x).view(np.uint32))
        return u ^ (0xFFFFFFFF if u >> 31 else 0x80000000)

    def score(c, q):
        x, y, gap = redf[c, 0], redf[c, 1], redf[c, 2]
        if q == 6:
            return fflip(-gap)
        if q > 6 and gap > 0:
            return None
        k = q - 7 if q > 6 else q
        return fflip(np.float32(x * cs[k]) + np.float32(y * sn[k]))

    want = set()
    for g, cs_ in groups.items():
        for q in range(13):
            cand = [(score(c, q), int(red[c, 5]), c) for c in cs_]
            cand = [t for t in cand if t[0] is not None]
            if cand:  # largest score, ties to the larger position identity; equal identities all win
                best = max(t[:2] for t in cand)
                want.update(t[2] for t in cand if t[:2] == best)
    kept = set(int(c) for c in np.nonzero(red[:, 6])[0])
    assert kept == want, (len(kept), len(want))
    if name == "tiles":
        assert len(kept) < 0.8 * len(listed)  # the reduction removes the inner points of the tiles' contacts
    if name == "tilted_tiles":  # groups mix touching and speculative contacts
        assert any(min(redf[c, 2] for c in cs_) <= 0 < max(redf[c, 2] for c in cs_) for cs_ in groups.values())
    # rows are the kept contacts (the coloring walks units in slot order, each unit's kept points together)
    keys = {(int(C[c][10]) | (int(C[c][11]) << 16)) for c in kept}
    assert len(info) == len(kept) and {int(info[r, 0]) for r in range(len(info))} == keys
    # group keys: the contact's two reduction groups, smaller first; a group's normals share one of 20 bins (the
    # Voronoi cells of a dodecahedron's vertices, angular radius 37.4 degrees)
    for g, cs_ in groups.items():
        for c in cs_:
            ga, gb = int(C[c][18]), int(C[c][19])
            assert (g[1], g[2]) == (min(ga, gb), max(ga, gb))
        N = np.array([C[c][4:7] * (-1 if int(C[c][19]) < int(C[c][18]) else 1) for c in cs_], np.float64)
        cosang = np.clip(N @ N.T, -1, 1)
        assert np.degrees(np.arccos(cosang.min())) <= 2 * 37.4 + 1e-3


# ---------------------------------------------------------------------------------------------------------------
# Generic manifold: invariants and a brute-force separating-axis search on random small hulls
# ---------------------------------------------------------------------------------------------------------------

def _hull_pair_scene(rng, dz=0.0, seed_pts=None):
    """Two random hulls of 6-12 points on spheres of radius 0.1, the second above the first at a random
    orientation and lowered by dz; no gravity, so nothing moves before the contact pass."""
    b = SceneBuilder(fps=240)
    pts, qa, qb, off = seed_pts
    ba = b.addBody(xform=[0, 0, 1.0, *qa], gravity=False)
    b.addShapeConvexHull(ba, vertices=pts[0].tolist(), gap=0.02)
    bb = b.addBody(xform=[off[0], off[1], 1.25 - dz, *qb], gravity=False)
    b.addShapeConvexHull(bb, vertices=pts[1].tolist(), gap=0.02)
    return b.finalize(solver=dict(substeps=1))


def _random_hull_pair(rng):
    pts = []
    for _ in range(2):
        n = int(rng.integers(6, 13))
        v = rng.normal(size=(n, 3))
        pts.append(0.1 * v / np.linalg.norm(v, axis=1, keepdims=True))
    return pts, _unit_quat(rng), _unit_quat(rng), rng.normal(size=2) * 0.03


def _world_hull(pw, o, s):
    """Vertices of shape s's hull in world coordinates: the packed hull vertices (scaled by the shape's scale)
    placed by the shape pose the substep's shape pass wrote (poses[P_SHAPE + 2s])."""
    SF, SU = pw.statF.reshape(-1, 4), pw.statU
    rec = SF[pw.h("W_SF_SHAPE") + s * 7:pw.h("W_SF_SHAPE") + s * 7 + 7]
    h = int(rec[0][1])
    v0, nv, f0, nf = (int(x) for x in SU[pw.h("W_SU_HULL") + h * 4:pw.h("W_SU_HULL") + h * 4 + 4])
    V = SF[pw.h("W_SF_VERTS") + v0:pw.h("W_SF_VERTS") + v0 + nv, :3].astype(np.float64) * rec[6][1:4].astype(np.float64)
    P = o.poses.reshape(-1, 4)
    ps = P[pw.h("W_P_SHAPE") + 2 * s, :3].astype(np.float64)
    qs = P[pw.h("W_P_SHAPE") + 2 * s + 1].astype(np.float64)
    return V @ _quat_mat(qs).T + ps


def _sat_separation(VA, VB):
    """Largest separation over face normals of both hulls' convex hulls and all edge-pair cross products (FP64,
    scipy hull as the independent face and edge source)."""
    from scipy.spatial import ConvexHull
    best = -1e30
    axes = []
    edges = []
    for V in (VA, VB):
        ch = ConvexHull(V)
        for eq in ch.equations:
            axes.append(eq[:3])
        E = set()
        for simp in ch.simplices:
            for i in range(3):
                a, c = sorted((simp[i], simp[(i + 1) % 3]))
                E.add((a, c))
        edges.append([V[c] - V[a] for a, c in E])
    for ea in edges[0]:
        for eb in edges[1]:
            n = np.cross(ea, eb)
            if np.linalg.norm(n) > 1e-9:
                axes.append(n / np.linalg.norm(n))
    for n in axes:
        s1 = (VB @ n).min() - (VA @ n).max()
        s2 = (VA @ n).min() - (VB @ n).max()
        best = max(best, s1, s2)
    return best


def _signed_dist_to_hull(V, p):
    from scipy.spatial import ConvexHull
    ch = ConvexHull(V)
    return float(np.max(ch.equations[:, :3] @ p + ch.equations[:, 3]))


@pytest.mark.parametrize("seed", range(12))
def test_generic_manifold_invariants_and_deepest_point_on_random_hulls(seed):
    """Hull-hull manifold (wgsl.js:875-1022): at most 5 points per pair, unit normal, every phi at most the pair's
    threshold, every point within the gap-inflated hulls, and the deepest phi equal to minus the penetration
    depth found by a brute-force separating-axis search over both hulls' faces and edge pairs."""
    rng = np.random.default_rng(100 + seed)
    depth = float(rng.uniform(0.002, 0.02))
    spec = _random_hull_pair(rng)
    dz = 0.0
    for _ in range(3):  # lower the upper hull until the brute-force separation is -depth
        pw = pack_scene(_hull_pair_scene(rng, dz, spec))
        o = OracleWorld(pw, f64=True)
        o.substeps(1, debug=True)
        WA, WB = _world_hull(pw, o, 0), _world_hull(pw, o, 1)
        sep = _sat_separation(WA, WB)
        dz += sep + depth
    assert -0.03 < sep < 0
    nc = int(o.dbg_bufs["ncontacts"][0])
    assert 1 <= nc <= 5
    C = o.dbg_bufs["contacts"][:nc * 24].reshape(nc, 24).astype(np.float64)
    _check_manifold(C, WA, WB, sep)


def _check_manifold(C, WA, WB, sep, thresh=0.04):
    for c in C:
        n = c[4:7]
        assert abs(np.linalg.norm(n) - 1) < 1e-6
        assert c[3] <= thresh + 1e-6
        for W in (WA, WB):
            d = _signed_dist_to_hull(W, c[0:3])
            assert d <= thresh + 1e-5
            if c[3] < 0:  # a penetrating point is the midpoint of a point of one hull and its partner |phi| along n
                assert d <= 0.5 * abs(c[3]) + 1e-6, (d, c[3])
    assert abs(C[:, 3].min() - sep) <= 1e-5, (C[:, 3].min(), sep)


def _flat_hull(rng, side):
    """6-7 points of a circle on the plane z = 0.05 * side plus 8 points of the sphere of radius 0.1 on the other
    side of that plane: a hull with one flat face, outward normal +z * side."""
    zc, rc = 0.05 * side, math.sqrt(0.1 ** 2 - 0.05 ** 2)
    k = int(rng.integers(6, 8))
    ph = rng.uniform(0, 2 * math.pi)
    pts = [[rc * math.cos(ph + 2 * math.pi * i / k), rc * math.sin(ph + 2 * math.pi * i / k), zc] for i in range(k)]
    while len(pts) < k + 8:
        v = rng.normal(size=3)
        v = 0.1 * v / np.linalg.norm(v)
        if v[2] * side < 0.03:
            pts.append(list(v))
    return np.array(pts)


@pytest.mark.parametrize("seed", range(8))
def test_generic_manifold_face_on_face_clips_the_overlap_polygon(seed):
    """Two hulls meeting flat face on parallel flat face (rotated about the normal, offset sideways): the manifold
    takes the clipping path (wgsl.js:875-937): 3 to 5 points, every one at the penetration depth found by the
    separating-axis search, inside both hulls, with the invariants above."""
    rng = np.random.default_rng(300 + seed)
    depth = float(rng.uniform(0.002, 0.01))
    pa, pb = _flat_hull(rng, +1), _flat_hull(rng, -1)
    aa, ab = rng.uniform(0, 2 * math.pi, size=2)
    off = rng.normal(size=2) * 0.02
    b = SceneBuilder(fps=240)
    ba = b.addBody(xform=[0, 0, 1.0, 0, 0, math.sin(aa / 2), math.cos(aa / 2)], gravity=False)
    b.addShapeConvexHull(ba, vertices=pa.tolist(), gap=0.02)
    bb = b.addBody(xform=[off[0], off[1], 1.1 - depth, 0, 0, math.sin(ab / 2), math.cos(ab / 2)], gravity=False)
    b.addShapeConvexHull(bb, vertices=pb.tolist(), gap=0.02)
    pw = pack_scene(b.finalize(solver=dict(substeps=1)))
    o = OracleWorld(pw, f64=True)
    o.substeps(1, debug=True)
    WA, WB = _world_hull(pw, o, 0), _world_hull(pw, o, 1)
    sep = _sat_separation(WA, WB)
    nc = int(o.dbg_bufs["ncontacts"][0])
    assert 3 <= nc <= 5, nc
    C = o.dbg_bufs["contacts"][:nc * 24].reshape(nc, 24).astype(np.float64)
    _check_manifold(C, WA, WB, sep)
    np.testing.assert_allclose(C[:, 3], sep, atol=1e-5)
    np.testing.assert_allclose(np.abs(C[:, 6]), 1.0, atol=1e-6)  # normal along z


# ---------------------------------------------------------------------------------------------------------------
# Friction patches (anchors): static friction holds a box on an incline, and does not below the friction angle
# ---------------------------------------------------------------------------------------------------------------

def _incline_box(theta, mu, anchors=True):
    b = SceneBuilder(fps=240)
    c, s_ = math.cos(theta / 2), math.sin(theta / 2)
    q = [s_, 0, 0, c]
    b.addShapeBox(-1, hx=3, hy=3, hz=0.1, gap=0.01, mu=mu, xform=[0, 0, 0, *q])
    n = np.array([0, -math.sin(theta), math.cos(theta)])
    p = n * (0.1 + 0.1 - 0.0005)
    box = b.addBody(xform=[*p, *q])
    b.addShapeBox(box, hx=0.1, hy=0.1, hz=0.1, gap=0.01, mu=mu)
    scene = b.finalize(solver=dict(substeps=4, iterations=8, **({} if anchors else dict(friction_anchor_beta=None))))
    return pack_scene(scene), p


def test_friction_patch_anchors_hold_a_box_on_an_incline_for_500_frames():
    """Friction patches with anchors (wgsl.js:1562-1725, 3829-3878): a box on a 20 degree slope with mu = 0.6
    (friction angle 31 degrees) stays where it settled within 0.1 mm over 500 frames."""
    pw, p0 = _incline_box(math.radians(20), 0.6)
    o = OracleWorld(pw)
    o.frames(20)
    start = o.poses[:3].astype(np.float64).copy()
    o.frames(500)
    drift = np.linalg.norm(o.poses[:3].astype(np.float64) - start)
    assert drift < 1e-4, drift
    assert np.abs(o.vel[:8]).max() < 1e-3


def test_friction_patch_lets_a_box_slide_above_the_friction_angle():
    """Above the friction angle (35 degrees, mu = 0.5, friction angle 26.6 degrees) the pooled friction load caps
    the patch: the box slides with the Coulomb acceleration g (sin theta - mu cos theta) within 3%."""
    theta, mu = math.radians(35), 0.5
    pw, p0 = _incline_box(theta, mu)
    o = OracleWorld(pw)
    o.frames(30)
    down = np.array([0, -math.cos(theta), -math.sin(theta)])
    v0 = float(o.vel[:3].astype(np.float64) @ down)
    o.frames(60)
    v1 = float(o.vel[:3].astype(np.float64) @ down)
    a = (v1 - v0) / (60 / 240)
    g = 9.81
    want = g * (math.sin(theta) - mu * math.cos(theta))
    assert abs(a - want) <= 0.03 * want, (a, want)
