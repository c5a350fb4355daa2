"""Scene -> packed world arrays (the host side of `Sim.build`, packages/steel-webgpu/src/sim.js:95-611).

The output is one `PackedWorld`: statF/statU with the reference's record layouts, the initial
poses/vel/art state, and the header of layout constants and solver parameters declared in
include/steel_world.h.  The step kernels and the CPU oracle read exactly these arrays.

Supported: the default option path and the off-default features every shipped scene uses (warm start,
sleeping, Nesterov momentum, velocity iterations, the four friction modes, contact compliance and torsion,
spin/rolling rows with the elliptic articulated cone, drive rows, implicit drives and springs, joint and body
velocity limits, joint friction, recorded target paths, the fader controller, kinematic bodies), FREE /
REVOLUTE / PRISMATIC / BALL / FIXED / D6 joints in tree articulations, and ball or revolute loop closures and
mimic joints.  Anything else (the pyramidal cone, free-rigid torsional/rolling overrides, cone shapes, shared
anchors, friction position iterations, row-schedule variants, other controllers) raises UnsupportedError, as the
reference raises for options it does not take (sim.js:73, 128, 159, 166, 376, 379, 411-420).
"""
from __future__ import annotations

import math
import re
import struct

import numpy as np

from .scene import JT, GT

LB = 268
L_V = 147
AVN = 14
MAX_SUBSTEPS = 32
COMBINE = ["average", "min", "multiply", "max"]


def _parse_fields():
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    txt = open(os.path.join(here, "..", "include", "steel_world.h")).read()
    body = txt[txt.index("enum steel_world_field"):txt.index("W_FIELD_COUNT")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b(W_[A-Z0-9_]+)\b", body)
    return {n: i for i, n in enumerate(names)}


FIELDS = _parse_fields()
HEADER_WORDS = 160


def _f2u(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def quat_rotate(q, v):
    """sim.js:16-22 (double precision)."""
    x, y, z, w = q
    ux = 2 * (x * v[0] + y * v[1] + z * v[2])
    s2 = w * w - (x * x + y * y + z * z)
    cx, cy, cz = y * v[2] - z * v[1], z * v[0] - x * v[2], x * v[1] - y * v[0]
    return [ux * x + s2 * v[0] + 2 * w * cx, ux * y + s2 * v[1] + 2 * w * cy, ux * z + s2 * v[2] + 2 * w * cz]


def inv3(m):
    """sim.js:24-31."""
    a, b, c, d, e, f, g, h, i = m
    A = e * i - f * h
    B = -(d * i - f * g)
    C = d * h - e * g
    det = a * A + b * B + c * C
    if abs(det) < 1e-30:
        return [0.0] * 9
    k = 1 / det
    return [A * k, -(b * i - c * h) * k, (b * f - c * e) * k, B * k, (a * i - c * g) * k, -(a * f - c * d) * k,
            C * k, -(a * h - b * g) * k, (a * e - b * d) * k]


def _next_pow2(n):
    p = 1
    while p < n:
        p <<= 1
    return p


class UnsupportedError(Exception):
    pass


class PackedWorld:
    """Static arrays, initial state and header of one world."""

    def __init__(self):
        self.header = np.zeros(HEADER_WORDS, np.uint32)

    def h(self, name):
        return int(self.header[FIELDS[name]])

    def hf(self, name):
        return struct.unpack("<f", struct.pack("<I", int(self.header[FIELDS[name]])))[0]


def capacities(NS, NB, NA, policy="reference"):
    """Per-world capacities (MAXP, MAXC, MAXU, UH, AH, MAXDENSE).

    "reference" keeps the reference's fixed sizes (sim.js:396, 439) wherever a world can reach them and
    shrinks each to the exact bound where it cannot (pairs <= NS(NS-1)/2, contacts <= 5 per pair).
    "compact" additionally bounds contacts by 16 per shape, for batches of many small worlds; a world
    that exceeds it reports the overflow in its stats.
    "scaled" keeps the reference's sizes as a floor and grows them with the world (24 pairs, 32 contacts and 12
    units per shape), for worlds whose resting contact graph outgrows the fixed sizes: 256 hulls settling in a
    2 m bin with the default 0.1 m gap reach about 3000 touching body pairs, and the units past the reference's
    2048 would get no rows, so their bodies sink through the floor (the reference reports that overflow in
    stats word 3).
    """
    maxp = max(1, min(16384, NS * (NS - 1) // 2))
    maxc = max(1, min(8192, 5 * maxp))
    if policy == "compact":
        maxc = max(1, min(maxc, max(256, 16 * NS)))
    if policy == "scaled":
        maxp = max(1, min(max(16384, 24 * NS), NS * (NS - 1) // 2))
        maxc = max(1, min(max(8192, 32 * NS), 5 * maxp))
        maxu = min(max(2048, 12 * NS), maxc)
        uh = min(max(4096, _next_pow2(2 * maxu)), maxc)
        ah = _next_pow2(max(64, 2 * maxu))
        maxdense = min(512, maxc) if NA > 0 else 1
        return maxp, maxc, maxu, uh, ah, maxdense
    maxu = min(2048, maxc)
    uh = min(4096, maxc)
    ah = min(4096, _next_pow2(max(64, 2 * maxu)))
    maxdense = min(512, maxc) if NA > 0 else 1
    return maxp, maxc, maxu, uh, ah, maxdense


# Shape count above which a world's listed collision filters are kept as sorted lists (W_FILT_LIST).
FILT_LIST_SHAPES = 16384
# Shape count above which the broadphase bins shapes in a uniform grid of cells (W_CELL) instead of sweeping on min x.
CELL_SHAPES = 4096


def cell_bound(statF, SF, NS):
    """Largest AABB extent the cell broadphase bins: a shape's AABB extent on any axis is at most
    2 (|hull half extents| + gap) whatever its rotation; planes and shapes over four times the median of that bound
    are left out (they are tested against every shape), so a few large static shapes do not coarsen the cells."""
    rec = statF[SF["SF_SHAPE"] * 4:(SF["SF_SHAPE"] + NS * 7) * 4].reshape(NS, 7, 4)
    hull = rec[:, 0, 1].astype(np.int64)
    hh = statF[SF["SF_HAABB"] * 4:].reshape(-1, 4)[2 * hull + 1, :3].astype(np.float64)
    b = 2.0 * (np.sqrt((hh * hh).sum(1)) + np.maximum(rec[:, 0, 3].astype(np.float64), 0.0))
    b = b[rec[:, 3, 0] != 4.0]
    if len(b) == 0:
        return 0.0
    b = b[b <= 4.0 * np.median(b)]
    return float(b.max()) if len(b) else 0.0

# Contact warm-start table slots per parity (sim.js:396 WH).
WARM_SLOTS = 16384


def sched_layout(MAXC, MAXU, UH, NCOL, NB, NCOLW, NL, WH=0, MAXDENSE=0):
    """Word offsets of the kernels' schedule tables (mirror of sched_layout in csrc/steel_dev.h)."""
    names = [("HKEY", 2 * UH), ("UKEY", 2 * MAXU), ("UTPW", 2 * MAXU), ("HCNT", UH), ("CSLOT", MAXC), ("CLOCAL", MAXC), ("USIZE", MAXU),
             ("USTART", MAXU + 1),
             ("CSTART", NCOL + 2), ("META", 8), ("CORDER", MAXC), ("RK0", MAXC), ("RK1", MAXC), ("RK2", MAXC), ("KEPT", MAXC),
             ("OSTART", MAXU), ("LEAD", MAXC), ("SLOT", MAXC * 26), ("ROWOF", MAXC), ("RUNIT", MAXC), ("PCOL", MAXC), ("PORD", MAXC),
             ("UTC", MAXU), ("UTP", MAXU), ("RCON", MAXC), ("UWOFF", UH), ("WALK", MAXC), ("HEADS", MAXC + 1), ("DSLOT", MAXC + 1),
             ("WCOL", MAXC), ("WMASK", max(NB, 1) * NCOLW), ("WCOUNT", NCOL + 2), ("LINKHIT", max(NB, 1)), ("JSPLIT", max(NL, 1)),
             ("SLOTT", MAXC * 4), ("NPAIRL", 4), ("UF", max(NB, 1) + max(NL, 1)), ("UFF", max(NB, 1) + max(NL, 1)),
             ("WWIN", 2 * WH), ("WSLOT", (MAXC + MAXDENSE) if WH else 0), ("FPX", MAXC * 4)]
    out, o = {}, 0
    for k, n in names:
        if k == "WWIN":
            o = (o + 1) & ~1  # uint64 entries
        out[k] = o
        o += n
    out["TOTAL"] = o
    return out


def world_sched_layout(pw):
    return sched_layout(pw.h("W_MAXC"), pw.h("W_MAXU"), pw.h("W_UH"), pw.h("W_NCOLW") * 32, pw.NB, pw.h("W_NCOLW"), pw.NL,
                        pw.h("W_WH"), pw.h("W_MAXDENSE"))


def pack_scene(scene, opts=None, capacity_policy="reference", capacity_override=None):
    """Pack a LoadedScene into a PackedWorld (sim.js:64-75 options, :95-611 build).

    capacity_override: {"MAXP"|"MAXC"|"MAXU"|"UH"|"AH"|"MAXDENSE": n} to force small capacities (overflow tests)."""
    S = scene.json
    f32, u32 = scene.f32, scene.u32
    opts = dict(opts or {})
    given = {k: v for k, v in opts.items() if v is not None}
    o = dict(substeps=S["substeps"], iterations=S["iterations"],
             regularization=(S.get("solver") or {}).get("contact_regularization", 0.0))
    o.update(given)
    if o["substeps"] > MAX_SUBSTEPS:
        raise ValueError(f"at most {MAX_SUBSTEPS} substeps")
    NB = S["bodies"]["count"]
    J = S["joints"]
    jointOfBody = [-1] * NB
    for j in range(J["count"]):
        jointOfBody[J["child"][j]] = j
    artJoints = {}
    for j in range(J["count"]):
        artJoints.setdefault(J["articulation"][j], []).append(j)
    bodyKind = [1] * NB
    bodyArt = [-1] * NB
    bodyLink = [-1] * NB
    arts, links, dofs = [], [], []
    SO = dict(S.get("solver") or {})
    SO.update(o.get("solver") or {})
    limitsOn = bool(SO.get("enable_joint_limits", False))
    vlimOn = bool(SO.get("enable_joint_velocity_limits", False))
    q0, tq0, tqd0 = [], [], []
    qMap, dofMap = {}, {}
    loops = []
    kin = S["bodies"].get("kinematic") or [False] * NB
    for art, js in artJoints.items():
        if art < 0:
            loops += js
            continue
        if len(js) == 1 and J["type"][js[0]] in (JT["FREE"], JT["DISTANCE"]):
            continue
        a = len(arts)
        linkStart, dofStart = len(links), len(dofs)
        localLink = {}
        for j in js:
            t = JT["FREE"] if J["type"][j] == JT["DISTANCE"] else J["type"][j]
            b = J["child"][j]
            l = len(links)
            localLink[b] = l
            parentLink = localLink[J["parent"][j]] if J["parent"][j] >= 0 else -1
            if t not in (JT["FIXED"], JT["REVOLUTE"], JT["PRISMATIC"], JT["BALL"], JT["FREE"], JT["D6"]):
                raise UnsupportedError(f"joint type {t} is not supported yet")
            ndof = J["qd_start"][j + 1] - J["qd_start"][j]
            nq = J["q_start"][j + 1] - J["q_start"][j]
            dof, qs = len(dofs), len(q0)
            scalar = t in (JT["REVOLUTE"], JT["PRISMATIC"], JT["D6"])
            for k in range(ndof):
                qd = J["qd_start"][j] + k
                limited = limitsOn and scalar and J["lower"][qd] > -1e5 and J["upper"][qd] < 1e5
                friction = J["friction"][qd] if SO.get("enable_joint_friction") else 0
                dofs.append(dict(armature=1e9 if kin[J["child"][j]] else J["armature"][qd], ke=J["ke"][qd], kd=J["kd"][qd],
                                 effort=J["effort"][qd], lower=J["lower"][qd], upper=J["upper"][qd], limited=1 if limited else 0,
                                 link=len(links), damping=J["passive_damping"][qd] if J.get("passive_damping") else 0,
                                 friction=friction, axis=J["axis"][qd], scalar=scalar,
                                 springKe=(J["spring_stiffness"][qd] if J.get("spring_stiffness") else 0) if scalar else 0,
                                 springRef=(J["spring_ref"][qd] if J.get("spring_ref") else 0),
                                 loadCoef=(J["friction_load_coefficient"][qd] or 0) if SO.get("enable_joint_friction") and scalar
                                 and J.get("friction_load_coefficient") else 0,
                                 velLimit=J["vel_limit"][qd] if scalar and vlimOn and J["vel_limit"][qd] < 1e5 else -1))
                tqd0.append(J["target_qd"][qd] if J.get("target_qd") else 0)
            for k in range(nq):
                qMap[J["q_start"][j] + k] = len(q0)
                q0.append(J["q"][J["q_start"][j] + k])
            for k in range(ndof):
                dofMap[J["qd_start"][j] + k] = dof + k
            src = J.get("target_q") or J["q"]
            for k in range(ndof):
                tq0.append(src[J["q_start"][j] + k] if scalar else 0)
            parentMovable = parentLink >= 0 and links[parentLink]["movable"]
            dd = J.get("dof_dim")
            links.append(dict(type=t, parent=parentLink, dof=dof, ndof=ndof, nlin=(dd[j][0] if dd else 0), qs=qs, body=b, joint=j,
                              movable=ndof > 0 or parentMovable))
            bodyKind[b] = 2
            bodyArt[b] = a
            bodyLink[b] = l
        arts.append(dict(linkStart=linkStart, linkCount=len(links) - linkStart, dofStart=dofStart, dofCount=len(dofs) - dofStart))
    NA, NL, ND = len(arts), len(links), len(dofs)
    if NA > 256:
        raise UnsupportedError(f"{NA} articulations; one solve workgroup handles at most 256")
    # Mimic joints (sim.js:160-172): follower q = c0 + c1 * leader q, per dof, as dense bilateral rows.
    mimics = []

    def add_mimic(f, r, c0, c1):
        for k in range(J["qd_start"][f + 1] - J["qd_start"][f]):
            fd, ld = dofMap.get(J["qd_start"][f] + k), dofMap.get(J["qd_start"][r] + k)
            a = bodyArt[J["child"][f]]
            if fd is None or ld is None or a != bodyArt[J["child"][r]]:
                raise ValueError("mimic joints must share an articulation")
            d0 = arts[a]["dofStart"]
            mimics.append(dict(art=a, f=fd - d0, l=ld - d0, c0=c0, c1=c1, fq=qMap[J["q_start"][f] + k], lq=qMap[J["q_start"][r] + k]))
    for f, r in enumerate(J.get("mimic_joint") or []):
        if r >= 0:
            add_mimic(f, r, J["mimic_coeffs"][f][0], J["mimic_coeffs"][f][1])
    for m in S.get("mimics") or []:
        if m.get("enabled"):
            add_mimic(m["follower"], m["leader"], m["c0"], m["c1"])
    ctrl = S.get("control")
    if ctrl and ctrl.get("type") != "fader":
        raise UnsupportedError(f"scene controller {ctrl.get('type')} is not supported")

    # Hulls (dedup by their blob range), unique edges with their two faces (sim.js:180-209).
    shapes = [s for s in S["shapes"] if s.get("hull") or s.get("plane")]
    hullIndex, hulls = {}, []
    for s in shapes:
        if not s.get("hull"):
            continue
        key = s["hull"]["verts"]["off"]
        if key not in hullIndex:
            hullIndex[key] = len(hulls)
            hulls.append(s["hull"])
    NS, NH = len(shapes), len(hulls)
    nVerts = sum(h["nv"] for h in hulls)
    nFaces = sum(h["nf"] for h in hulls)
    nFv = sum(h["face_verts"]["n"] for h in hulls)
    hullEdges = []
    for h in hulls:
        fs, fv = u32(h["face_start"]), u32(h["face_verts"])
        faces, edges = {}, []
        for f in range(h["nf"]):
            n = int(fs[f + 1] - fs[f])
            for i in range(n):
                a, b = int(fv[fs[f] + i]), int(fv[fs[f] + (i + 1) % n])
                key = min(a, b) * 65536 + max(a, b)
                faces.setdefault(key, []).append(f)
                if a < b:
                    edges.append((a, b, key))
        hullEdges.append([(a, b, faces[key][0], faces[key][-1]) for a, b, key in edges])
    nEdges = sum(len(e) for e in hullEdges)

    for s in shapes:
        if s.get("plane"):
            continue
        pk = (s.get("prim") or {}).get("kind", 0)
        t = s["type"]
        if t == GT["CONE"]:
            raise UnsupportedError("cone shapes are not supported yet")

    SF = {}
    off = 0
    SF["SF_BODY"] = off; off += NB * 8
    SF["SF_SHAPE"] = off; off += NS * 7
    SF["SF_HAABB"] = off; off += NH * 2
    SF["SF_VERTS"] = off; off += nVerts
    SF["SF_FPLANE"] = off; off += nFaces
    SF["SF_LINK"] = off; off += max(NL, 1) * 6
    SF["SF_DOF"] = off; off += max(ND, 1) * 5
    pathFloats = (S.get("target_frames") or 0) * (S.get("driven") or 0)
    SF["SF_PATH"] = off; off += (pathFloats + 3) // 4
    SF["SF_MIMIC"] = off; off += max(len(mimics), 1) * 2
    SF["SF_LOOP"] = off; off += max(len(loops), 1) * 5
    SF["SF_CTRL"] = off; off += 1
    nstatf = off
    statF = np.zeros(nstatf * 4, np.float32)
    SU = {}
    uoff = 0
    SU["SU_HULL"] = uoff; uoff += NH * 4
    SU["SU_FACE"] = uoff; uoff += nFaces * 2
    SU["SU_FV"] = uoff; uoff += nFv
    SU["SU_HEDGE"] = uoff; uoff += NH * 2
    SU["SU_EDGE"] = uoff; uoff += nEdges * 4
    SU["SU_ART"] = uoff; uoff += max(NA, 1) * 4
    SU["SU_ACON"] = uoff; uoff += max(NA, 1)

    def groupsMeet(ga, gb):
        return ga != 0 and gb != 0 and ((ga == gb or gb < 0) if ga > 0 else ga != gb)

    index = {}
    for i, sh in enumerate(S["shapes"]):
        if sh.get("hull") or sh.get("plane"):
            index[i] = len(index)
    fl = [(index.get(a0), index.get(b0)) for a0, b0 in S.get("filters", [])]
    fl = [(a, b) for a, b in fl if a is not None and b is not None and a != b]
    # World and group rules (sim.js:244-255).  Both are symmetric in the two shapes, so they are decided once per
    # pair of (world, group) classes; a class pair that does not meet filters all of its shape pairs.
    cls = {}
    for a in range(NS):
        wa, ga = shapes[a].get("world", -1), shapes[a].get("group", -1)
        cls.setdefault((-1 if wa is None else wa, -1 if ga is None else ga), []).append(a)
    keys = sorted(cls)
    apart = [(x, y) for x in keys for y in keys if x <= y and not ((x[0] < 0 or y[0] < 0 or x[0] == y[0]) and groupsMeet(x[1], y[1]))]
    # Past FILT_LIST_SHAPES shapes, a world whose filters are only listed pairs keeps them as per-shape sorted lists
    # instead of the NS x NS bitset (about 1.3 GB at 100k shapes).
    filt_list = NS > FILT_LIST_SHAPES and not apart
    if filt_list:
        lists = [[] for _ in range(NS)]
        for a, b in sorted({(min(a, b), max(a, b)) for a, b in fl}):
            lists[a].append(b)
        offs = np.zeros(NS + 1, np.int64)
        offs[1:] = np.cumsum([len(x) for x in lists])
        SU["SU_FILTER"] = uoff; uoff += NS + 1 + int(offs[-1])
    else:
        SU["SU_FILTER"] = uoff; uoff += (NS * NS + 31) // 32
    nstatu = uoff
    statU = np.zeros(max(nstatu, 1), np.uint32)
    filt = statU[SU["SU_FILTER"]:]
    if filt_list:
        filt[:NS + 1] = offs
        filt[NS + 1:NS + 1 + int(offs[-1])] = [j for x in lists for j in x]

    def setbits(a, b):  # pairs (a, b) and (b, a); bit index a * NS + b needs 64 bits past 65536 shapes
        a, b = np.asarray(a, np.int64), np.asarray(b, np.int64)
        for k in (a * NS + b, b * NS + a):
            np.bitwise_or.at(filt, k >> 5, (np.uint32(1) << (k & 31).astype(np.uint32)))

    if not filt_list:
        if fl:
            setbits([a for a, _ in fl], [b for _, b in fl])
        for x, y in apart:
            A, B = np.asarray(cls[x], np.int64), np.asarray(cls[y], np.int64)
            aa, bb = np.meshgrid(A, B, indexing="ij")
            aa, bb = aa.ravel(), bb.ravel()
            keep = aa != bb
            setbits(aa[keep], bb[keep])

    # Bodies (sim.js:257-305).
    bodyPose = np.zeros(NB * 8, np.float32)
    pw = PackedWorld()
    pw.bodyKind = list(bodyKind)
    # Bodies the mouse spring may pull (sim.js:175-178): free bodies and movable links of undriven articulations.
    kinem = (S["bodies"].get("kinematic") or [0] * NB)
    artDriven = [any(d["ke"] > 0 or d["kd"] > 0 for d in dofs[a["dofStart"]:a["dofStart"] + a["dofCount"]]) for a in arts]
    pw.draggable = [S.get("mouse_drag") is not False and not kinem[b] and
                    (k == 1 or (k == 2 and links[bodyLink[b]]["movable"] and not artDriven[bodyArt[b]])) for b, k in enumerate(bodyKind)]
    for b in range(NB):
        o4 = (SF["SF_BODY"] + b * 8) * 4
        m = S["bodies"]["mass"][b]
        link = bodyLink[b]
        kind = 3 if (bodyKind[b] == 2 and not links[link]["movable"]) else bodyKind[b]
        pw.bodyKind[b] = kind
        statF[o4] = 1 / m if m > 0 and not kin[b] else 0  # kinematic: infinite mass (sim.js:266-275)
        statF[o4 + 1] = kind
        statF[o4 + 2] = bodyArt[b]
        statF[o4 + 3] = link
        com = S["bodies"]["com"][b]
        statF[o4 + 4:o4 + 8] = [com[0], com[1], com[2], max(m, 0)]
        I = S["bodies"]["inertia"][b]
        Ii = inv3(I)
        for r in range(3):
            statF[o4 + 8 + r * 4:o4 + 11 + r * 4] = I[r * 3:r * 3 + 3]
            statF[o4 + 20 + r * 4:o4 + 23 + r * 4] = [0, 0, 0] if kin[b] else Ii[r * 3:r * 3 + 3]
        bq = S["bodies"]["q"][b]
        q = [bq[3], bq[4], bq[5], bq[6]]
        c = quat_rotate(q, com)
        bodyPose[b * 8:b * 8 + 8] = [bq[0] + c[0], bq[1] + c[1], bq[2] + c[2], 0] + q
    radius = [0.0] * NB
    for s in shapes:
        if s["body"] < 0 or s.get("plane"):
            continue
        com = S["bodies"]["com"][s["body"]]
        pr = s.get("prim") or {}
        sr = pr["r"] if pr.get("kind") == 1 else (pr["r"] + pr["h"] if pr.get("kind") == 2 else 0.0)
        if s.get("hull") and not pr.get("kind"):
            v = f32(s["hull"]["verts"]).astype(np.float64)
            for i in range(s["hull"]["nv"]):
                sr = max(sr, math.hypot(v[i * 3], v[i * 3 + 1], v[i * 3 + 2]))
        d = math.hypot(s["xf"][0] - com[0], s["xf"][1] - com[1], s["xf"][2] - com[2])
        radius[s["body"]] = max(radius[s["body"]], d + sr)
    for b in range(NB):
        o4 = (SF["SF_BODY"] + b * 8) * 4
        statF[o4 + 11] = radius[b]
        ng = S["bodies"].get("no_gravity") or [False] * NB
        statF[o4 + 31] = (1 if kin[b] else 0) | (2 if ng[b] else 0)
        lim = lambda k: -1 if (S["bodies"].get(k) or [None] * NB)[b] is None else S["bodies"][k][b]
        statF[o4 + 15] = (S["bodies"].get("angular_damping") or [0.05] * NB)[b]
        statF[o4 + 19] = lim("max_linear_velocity")
        statF[o4 + 23] = lim("max_angular_velocity")
        statF[o4 + 27] = lim("max_depenetration_velocity")
    # the reference switches the free-body limit pass on when either array is present (sim.js:471)
    vlimFree = S["bodies"].get("max_linear_velocity") is not None or S["bodies"].get("max_angular_velocity") is not None
    # Reduction groups (sim.js:308-313).
    rkeys = {}
    rgroup = []
    for i, s in enumerate(shapes):
        k = f"s{i}" if s["body"] < 0 else (s["body"], s["mu"], s.get("restitution"), s.get("mu_torsional"), s.get("mu_rolling"),
                                           s.get("ke"), s.get("kd"), s.get("condim"), s.get("torsion_radius"))
        if k not in rkeys:
            rkeys[k] = len(rkeys)
        rgroup.append(rkeys[k])
    GTN = S["geo_type_names"]

    def geoScale(s):
        if s["type"] in (GTN["SPHERE"], GTN["CAPSULE"]):
            return [s["prim"]["r"], s["prim"]["h"], 0]
        if s["type"] not in (GTN["BOX"], GTN["CYLINDER"]):
            return [1, 1, 1]
        v = f32(s["hull"]["verts"])
        m = [0.0, 0.0, 0.0]
        for i in range(len(v)):
            m[i % 3] = max(m[i % 3], abs(float(v[i])))
        return m if s["type"] == GTN["BOX"] else [m[0], m[2], 0]

    for i, s in enumerate(shapes):
        o4 = (SF["SF_SHAPE"] + i * 7) * 4
        pr = s.get("prim") or {}
        statF[o4:o4 + 4] = [s["body"], hullIndex[s["hull"]["verts"]["off"]] if s.get("hull") else 0, s["mu"],
                            s["gap"] if s.get("gap") is not None else 0.001]
        statF[o4 + 4:o4 + 12] = [s["xf"][0], s["xf"][1], s["xf"][2], 0, s["xf"][3], s["xf"][4], s["xf"][5], s["xf"][6]]
        statF[o4 + 20:o4 + 24] = [s.get("ke") or 0, s.get("kd") or 0, rgroup[i], 0]
        statF[o4 + 12:o4 + 16] = [4 if s.get("plane") else pr.get("kind", 0), pr.get("r", 0), pr.get("h", 0), s.get("torsion_radius") or 0]
        statF[o4 + 16:o4 + 20] = [s.get("restitution") or 0, s.get("mu_torsional") or 0, s.get("mu_rolling") or 0,
                                  s.get("condim") if s.get("condim") is not None else 3]
        statF[o4 + 24:o4 + 28] = [GTN["PLANE"] if s.get("plane") else s["type"]] + geoScale(s)
    # Hull geometry (sim.js:335-355).
    vOff = fOff = fvOff = eOff = 0
    for hi, h in enumerate(hulls):
        statU[SU["SU_HEDGE"] + hi * 2:SU["SU_HEDGE"] + hi * 2 + 2] = [eOff, len(hullEdges[hi])]
        for k, e in enumerate(hullEdges[hi]):
            statU[SU["SU_EDGE"] + (eOff + k) * 4:SU["SU_EDGE"] + (eOff + k) * 4 + 4] = e
        eOff += len(hullEdges[hi])
        v = f32(h["verts"])
        planes = f32(h["face_planes"])
        fs = u32(h["face_start"])
        fv = u32(h["face_verts"])
        lo = [1e30, 1e30, 1e30]
        hiB = [-1e30, -1e30, -1e30]
        for i in range(h["nv"]):
            for k in range(3):
                lo[k] = min(lo[k], float(v[i * 3 + k]))
                hiB[k] = max(hiB[k], float(v[i * 3 + k]))
            statF[(SF["SF_VERTS"] + vOff + i) * 4:(SF["SF_VERTS"] + vOff + i) * 4 + 4] = [v[i * 3], v[i * 3 + 1], v[i * 3 + 2], 0]
        base = (SF["SF_HAABB"] + hi * 2) * 4
        statF[base:base + 8] = [(lo[0] + hiB[0]) / 2, (lo[1] + hiB[1]) / 2, (lo[2] + hiB[2]) / 2, 0,
                                (hiB[0] - lo[0]) / 2, (hiB[1] - lo[1]) / 2, (hiB[2] - lo[2]) / 2, 0]
        statU[SU["SU_HULL"] + hi * 4:SU["SU_HULL"] + hi * 4 + 4] = [vOff, h["nv"], fOff, h["nf"]]
        for f in range(h["nf"]):
            statF[(SF["SF_FPLANE"] + fOff + f) * 4:(SF["SF_FPLANE"] + fOff + f) * 4 + 4] = planes[f * 4:f * 4 + 4]
            statU[SU["SU_FACE"] + (fOff + f) * 2:SU["SU_FACE"] + (fOff + f) * 2 + 2] = [fvOff + int(fs[f]), int(fs[f + 1] - fs[f])]
        statU[SU["SU_FV"] + fvOff:SU["SU_FV"] + fvOff + len(fv)] = fv
        vOff += h["nv"]
        fOff += h["nf"]
        fvOff += len(fv)
    # Articulations (sim.js:357-368).
    for i, l in enumerate(links):
        o4 = (SF["SF_LINK"] + i * 6) * 4
        xp, xc = J["X_p"][l["joint"]], J["X_c"][l["joint"]]
        ax = J["axis"][J["qd_start"][l["joint"]]] if l["ndof"] == 1 else [0, 0, 1]
        statF[o4:o4 + 4] = [l["type"], l["parent"], l["dof"], l["body"]]
        statF[o4 + 4:o4 + 12] = [xp[0], xp[1], xp[2], l["ndof"], xp[3], xp[4], xp[5], xp[6]]
        statF[o4 + 12:o4 + 20] = [xc[0], xc[1], xc[2], l["qs"], xc[3], xc[4], xc[5], xc[6]]
        statF[o4 + 20:o4 + 24] = [ax[0], ax[1], ax[2], l["nlin"] if l["type"] == JT["D6"] else 0]
    for i, d in enumerate(dofs):
        o4 = (SF["SF_DOF"] + i * 5) * 4
        statF[o4:o4 + 20] = [d["armature"], d["ke"], d["kd"], d["effort"], d["lower"], d["upper"], d["limited"], d["link"],
                             d["damping"], d["friction"], d["velLimit"], 1 if d["scalar"] else 0, d["axis"][0], d["axis"][1], d["axis"][2],
                             d["loadCoef"], d["springKe"], d["springRef"], 0, 0]
    for i, a in enumerate(arts):
        statU[SU["SU_ART"] + i * 4:SU["SU_ART"] + i * 4 + 4] = [a["linkStart"], a["linkCount"], a["dofStart"], a["dofCount"]]

    # Solver options (sim.js:398-437).
    ncolw = o.get("colorWords", 2)
    SV = dict(body_pair_cell_size=0.06, cfm=1e-6, relaxation=1, baumgarte=0.2, contact_speculative_scale=1, contact_gap_gate=None,
              velocity_limit_activation_fraction=0, contact_friction_gap_max=None, contact_friction_scale=1, contact_shared_anchor=False,
              contact_friction_shared_anchor=False, contact_warmstart_decay=1, restitution_velocity_threshold=0.5, enable_restitution=True,
              contact_friction_position_iterations=-1, velocity_iterations=0, joint_limit_activation_gap=None)
    SV.update(S.get("solver") or {})
    SV.update(o.get("solver") or {})
    warm = o.get("warmStart")
    if warm is None:
        warm = 1 if SV.get("contact_warmstart") else 0
    fmodes = ["fixed_normal", "bisection", "bisection_desaxce", "coulomb_newton"]
    fm = SV.get("friction_mode") or "fixed_normal"
    if fm not in fmodes:
        raise ValueError(f"friction_mode {fm} is not supported")
    fmode = fmodes.index(fm)
    for k, default in (("enable_bilateral_preelimination", False), ("propagation_same_articulation_rows", False), ("update_mass_matrix_interval", 1)):
        if SV.get(k) is not None and SV.get(k) != default:
            raise ValueError(f"{k}={SV[k]} is not supported")
    if SV.get("torsional_rolling_friction_cone") == "pyramidal":
        raise UnsupportedError("the pyramidal torsional/rolling cone is not supported yet")
    for k in ("contact_shared_anchor", "contact_friction_shared_anchor"):
        if SV.get(k) or SO.get(k):
            raise UnsupportedError(f"{k} is not supported yet")
    if SV.get("contact_friction_position_iterations", -1) >= 0:
        raise UnsupportedError("contact_friction_position_iterations is not supported yet")
    if SV.get("row_schedule") == "contact_then_internal" or SV.get("propagation_color_order") == "longest_first":
        raise UnsupportedError("row_schedule/propagation_color_order variants are not supported yet")
    if SV.get("free_rigid_torsional_friction", -1) > 0 or SV.get("free_rigid_rolling_friction", -1) > 0:
        raise UnsupportedError("free-body torsional/rolling friction is not supported yet")
    compl = bool(SV.get("contact_compliance"))
    anchorBeta = SV.get("friction_anchor_beta", 0.2)
    angArt = bool(SV.get("enable_torsional_rolling_friction"))
    if anchorBeta is None or compl or fmode > 0 or angArt:  # these turn anchors off (sim.js:421-422)
        anchorBeta = 0.0
    angOn = angArt or any((s.get("condim", 3) or 3) >= 4 for s in S["shapes"])  # sim.js:425-429 (free-rigid overrides refused above)
    PO = 34 if angOn else 19
    FMO = PO + (4 if anchorBeta > 0 else 0)
    nest = bool(SV.get("nesterov_momentum"))
    tors = any((s.get("torsion_radius") or 0) > 0 for s in S["shapes"])
    TO = FMO + (2 if fmode > 0 else 0)
    EO = TO + (5 if tors else 0)
    CO = EO + (4 if angArt else 0)
    MO = CO + (1 if compl else 0)
    ROWS = MO + (2 if nest else 0)
    P = dict(P_BODY=0, P_SHAPE=NB * 2, P_AABB=NB * 2 + NS * 2, P_PREV=NB * 2 + NS * 4)
    maxp, maxc, maxu, uh, ah, maxdense = capacities(NS, NB, NA, capacity_policy)
    if capacity_override:
        co = capacity_override
        maxp, maxc = co.get("MAXP", maxp), co.get("MAXC", maxc)
        maxu, uh = min(co.get("MAXU", maxu), maxc), min(co.get("UH", uh), maxc)
        ah, maxdense = co.get("AH", ah), co.get("MAXDENSE", maxdense)
    NQ = len(q0)
    NDA = max([1] + [a["dofCount"] for a in arts])
    DHDR = 25
    DPS = DHDR + 6 * NDA
    A = {}
    ao = 0
    for k, n in (("A_Q", NQ), ("A_QD", ND), ("A_QDH", ND), ("A_INT", ND), ("A_TGT", ND), ("A_TGTV", ND), ("A_LIM", 2 * ND), ("A_DRV", ND),
                 ("A_JFL", ND), ("A_JF", ND), ("A_BF", NB * 8), ("A_SLP", (NB + NA) * 4), ("A_FRIC", ND), ("A_HINV", ND),
                 ("A_LINK", max(NL, 1) * LB), ("A_SNAP", NB * 8 + ND), ("A_VLG", ND + NB * 6), ("A_AVL", NA), ("A_LG", ND), ("A_RL", ND),
                 ("A_RLN", NA), ("A_ULOAD", maxu if anchorBeta > 0 else 0), ("A_DENSE", maxdense * DPS)):
        A[k] = ao
        ao += max(n, 1)
    depth = []
    for l in links:
        depth.append(depth[l["parent"]] + 1 if l["parent"] >= 0 else 1)
    MAXPATH = max([1] + depth)
    MAXLP = max([1] + [a["linkCount"] for a in arts])
    JM = 6 if any(l["ndof"] > 3 for l in links) else 3

    artInit = np.zeros(ao, np.float32)
    artInit[A["A_Q"]:A["A_Q"] + NQ] = q0
    artInit[A["A_TGT"]:A["A_TGT"] + len(tq0)] = tq0
    artInit[A["A_TGTV"]:A["A_TGTV"] + len(tqd0)] = tqd0
    bodyVel = np.zeros(NB * 8, np.float32)
    tw = S["bodies"].get("twist")
    if tw:
        for b, t in enumerate(tw):
            if bodyKind[b] == 1:
                bodyVel[b * 8:b * 8 + 8] = [t[0], t[1], t[2], 0, t[3], t[4], t[5], 0]
    for l in links:
        t = tw[l["body"]] if tw else None
        if l["type"] != JT["FREE"] or l["parent"] >= 0 or not t:
            continue
        bq = S["bodies"]["q"][l["body"]]
        rc = quat_rotate([bq[3], bq[4], bq[5], bq[6]], S["bodies"]["com"][l["body"]])
        w = [t[3], t[4], t[5]]
        v = [t[0] - (w[1] * rc[2] - w[2] * rc[1]), t[1] - (w[2] * rc[0] - w[0] * rc[2]), t[2] - (w[0] * rc[1] - w[1] * rc[0])]
        artInit[A["A_QD"] + l["dof"]:A["A_QD"] + l["dof"] + 6] = v + w

    fps = S["fps"]
    dt = 1 / fps / o["substeps"]
    grav = list(S["gravity"])
    hdr = pw.header
    H = lambda k, v: hdr.__setitem__(FIELDS[k], int(v))
    HF = lambda k, v: hdr.__setitem__(FIELDS[k], _f2u(v))
    for k, v in (("W_NB", NB), ("W_NS", NS), ("W_NH", NH), ("W_NA", NA), ("W_NL", NL), ("W_ND", ND), ("W_NQ", NQ), ("W_NDA", NDA),
                 ("W_JM", JM), ("W_MAXPATH", MAXPATH), ("W_MAXLP", MAXLP), ("W_MAXP", maxp), ("W_MAXC", maxc), ("W_MAXU", maxu),
                 ("W_UH", uh), ("W_AH", ah), ("W_MAXDENSE", maxdense), ("W_NCOLW", ncolw), ("W_ROWS", ROWS), ("W_PO", PO),
                 ("W_PATCH", 1 if anchorBeta > 0 else 0), ("W_EXTRA", 1 if anchorBeta > 0 else 0),
                 ("W_FRICTION_COMBINE", COMBINE.index(SO.get("friction_combine") or "average")),
                 ("W_RESTITUTION_COMBINE", COMBINE.index(SO.get("restitution_combine") or "average")),
                 ("W_REST", 1 if (SV.get("enable_restitution") and not SV.get("contact_compliance")) else 0),
                 ("W_NSTATF", nstatf), ("W_NSTATU", nstatu), ("W_NPOSES", NB * 4 + NS * 4), ("W_NART", ao),
                 ("W_DHDR", DHDR), ("W_DPS", DPS), ("W_NWARM", 2 * ah + 2 * ah * AVN),
                 ("W_ITERS", o["iterations"]), ("W_SUBSTEPS", o["substeps"])):
        H(k, v)
    for k, v in SF.items():
        H("W_" + k, v)
    for k, v in SU.items():
        H("W_" + k, v)
    for k, v in P.items():
        H("W_" + k, v)
    for k in ("A_Q", "A_QD", "A_QDH", "A_TGT", "A_TGTV", "A_LIM", "A_DRV", "A_JF", "A_BF", "A_FRIC", "A_HINV", "A_LINK", "A_LG", "A_RL",
              "A_RLN", "A_ULOAD", "A_DENSE", "A_VLG", "A_AVL", "A_JFL", "A_INT", "A_SLP", "A_SNAP"):
        H("W_" + k, A[k])
    H("W_VLIM_ON", 1 if vlimOn else 0)
    H("W_VLIM_FREE", 1 if vlimFree else 0)
    H("W_JFL", 1 if any(d["loadCoef"] > 0 for d in dofs) else 0)
    if pathFloats:
        statF[SF["SF_PATH"] * 4:SF["SF_PATH"] * 4 + pathFloats] = f32(S["targets"])[:pathFloats]
    H("W_DRIVEN", S.get("driven") or 0)
    H("W_PATH_FRAMES", S.get("target_frames") or 0)
    H("W_INTEGRAL_COUNT", S.get("integral_count") or 0)
    H("W_TGT_HELD", 1 if S.get("targets_held") else 0)
    H("W_SLEEP", 1 if (SV.get("enable_sleeping") or SO.get("enable_sleeping")) else 0)
    H("W_NMIMIC", len(mimics))
    H("W_FILT_LIST", 1 if filt_list else 0)
    HF("W_CELL", cell_bound(statF, SF, NS) if NS > CELL_SHAPES else 0.0)
    H("W_WH", WARM_SLOTS if warm else 0)
    H("W_NEST", 1 if nest else 0)
    H("W_MO", MO)
    H("W_TORS", 1 if tors else 0)
    H("W_VELITERS", int(SV.get("velocity_iterations") or 0))
    H("W_COMPL", 1 if compl else 0)
    H("W_DRIVE_ROWS", 1 if SV.get("drive_mode") == "rows" else 0)
    H("W_FUSE_VLIM", 0 if SV.get("fuse_joint_velocity_limits") is False else 1)
    H("W_VEL_DRIVE_ACTIVE", 1 if SV.get("velocity_drive_mode") == "active" else 0)
    H("W_CO", CO)
    H("W_TO", TO)
    for i, m in enumerate(mimics):
        statF[(SF["SF_MIMIC"] + i * 2) * 4:(SF["SF_MIMIC"] + i * 2) * 4 + 8] = [m["art"], m["f"], m["l"], m["c1"], m["fq"], m["lq"], m["c0"], 0]
        statU[SU["SU_ACON"] + m["art"]] = 1  # articulations owning a mimic never sleep (sim.js:369-370)
    # Loop closures (ball or revolute) inside one articulation or to the world (sim.js:372-385).
    H("W_NLOOP", len(loops))
    H("W_ANG", 1 if angOn else 0)
    H("W_FMODE", fmode)
    H("W_ANG_ART", 1 if angArt else 0)
    H("W_EO", EO)
    H("W_FMO", FMO)
    if ctrl:  # the fader controller (sim.js:388-393)
        H("W_CTRL_KIND", 1)
        H("W_CTRL_A", qMap[ctrl["fader_q"]])
        H("W_CTRL_B", dofMap[ctrl["wheel_v"]])
        statF[SF["SF_CTRL"] * 4:SF["SF_CTRL"] * 4 + 4] = [ctrl["travel"], ctrl["max_speed"], 0, 0]
    for i, j in enumerate(loops):
        t = J["type"][j]
        if t not in (JT["BALL"], JT["REVOLUTE"]):
            raise UnsupportedError(f"loop joint type {t} is not supported")
        world = J["parent"][j] < 0
        cb = J["child"][j]
        pl = -1 if world else bodyLink[J["parent"][j]]
        cl = bodyLink[cb]
        if cl < 0 or (not world and (pl < 0 or bodyArt[J["parent"][j]] != bodyArt[cb])):
            raise ValueError("loop joints must close inside one articulation or to the world")
        xp, xc = J["X_p"][j], J["X_c"][j]
        qd0 = J["qd_start"][j]
        ax = J["axis"][qd0] if qd0 < len(J["axis"]) else [0, 0, 1]
        ap, ac = quat_rotate([xp[3], xp[4], xp[5], xp[6]], ax), quat_rotate([xc[3], xc[4], xc[5], xc[6]], ax)
        statF[(SF["SF_LOOP"] + i * 5) * 4:(SF["SF_LOOP"] + i * 5 + 5) * 4] = [pl, cl, bodyArt[cb], 1 if t == JT["REVOLUTE"] else 0,
                                                                          xp[0], xp[1], xp[2], 0, xc[0], xc[1], xc[2], 0,
                                                                          *ap, 0, *ac, 0]
        statU[SU["SU_ACON"] + bodyArt[cb]] = 1  # articulations owning a loop closure never sleep
    gate = SV.get("contact_gap_gate")
    fgap = SV.get("contact_friction_gap_max")
    jlg = SV.get("joint_limit_activation_gap")
    for k, v in (("W_F_CFM", SV["cfm"]), ("W_F_OMEGA", SV["relaxation"]), ("W_F_SPEC", SV["contact_speculative_scale"]),
                 ("W_F_GAP_GATE", 1e30 if gate is None else gate), ("W_F_FGAP", 1e30 if fgap is None else fgap),
                 ("W_F_CREEP", SV.get("torsional_rolling_friction_creep_speed") or 0),
                 ("W_F_FSCALE", SV["contact_friction_scale"]), ("W_F_ANCHOR_BETA", anchorBeta), ("W_F_BP_CELL", SV["body_pair_cell_size"]),
                 ("W_F_REST_SPEED", SV["restitution_velocity_threshold"]), ("W_F_JL_GAP", min(0.05 if jlg is None else jlg, 1e30)),
                 ("W_F_JL_SPEED", 2 if jlg is None else 0),
                 ("W_DT", dt), ("W_BETA", SV["baumgarte"]), ("W_MAX_DEPEN", 2.0), ("W_GX", grav[0]), ("W_GY", grav[1]), ("W_GZ", grav[2]),
                 ("W_WARM_SCALE", warm * SV["contact_warmstart_decay"]), ("W_REG", 0 if SV.get("contact_compliance") else o["regularization"]),
                 ("W_FRAME_DT", 1 / fps), ("W_F_VLIM_FRAC", SV["velocity_limit_activation_fraction"]), ("W_FPS", fps),
                 ("W_F_SLEEP_LIN", SV.get("sleep_linear_threshold") if SV.get("sleep_linear_threshold") is not None else 0.05),
                 ("W_F_SLEEP_ANG", SV.get("sleep_angular_threshold") if SV.get("sleep_angular_threshold") is not None else 0.15),
                 ("W_F_SLEEP_TIME", SV.get("sleep_quiet_time") if SV.get("sleep_quiet_time") is not None else 0.5)):
        HF(k, v)

    pw.statF = statF
    pw.statU = statU
    pw.poses0 = np.zeros((NB * 4 + NS * 4) * 4, np.float32)
    pw.poses0[:NB * 8] = bodyPose
    pw.vel0 = bodyVel
    pw.art0 = artInit
    pw.warm0 = np.zeros(2 * ah + 2 * ah * AVN, np.uint32)
    pw.NB, pw.NS, pw.NA, pw.NL, pw.ND = NB, NS, NA, NL, ND
    pw.links = links
    pw.dofMap = dofMap
    pw.qMap = qMap
    pw.A = A
    pw.SF, pw.SU, pw.P = SF, SU, P
    pw.opts = o
    pw.solver = SV
    pw.fps = fps
    pw.scene = scene
    return pw
