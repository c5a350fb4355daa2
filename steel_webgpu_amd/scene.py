"""Scene construction and scene-folder I/O, mirroring the reference's JS host.

`SceneBuilder` follows packages/steel-webgpu/src/builder.js:241-490 (Newton ModelBuilder analogue):
the same method names, option names, defaults and the same `finalize()` output (a scene dict plus a
binary blob with float32/uint32 ranges), so scenes built here load in the reference and vice versa.
`load_scene` reads the reference's scene-folder format (scene.json + scene.bin,
packages/steel-webgpu/src/sim.js:40-61; field list tools/newton_export.py:85-315).

Convex hulls are computed with Qhull (scipy) and merged into planar polygons with CCW loops; the hull's
vertex/face order is an input of the step (it sets feature ids), so a scene carries its topology in
scene.bin and every consumer reads the same arrays.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

JT = {"PRISMATIC": 0, "REVOLUTE": 1, "BALL": 2, "FIXED": 3, "FREE": 4, "DISTANCE": 5, "D6": 6}
GT = {"NONE": 0, "PLANE": 1, "HFIELD": 2, "SPHERE": 3, "CAPSULE": 4, "ELLIPSOID": 5, "CYLINDER": 6,
      "BOX": 7, "MESH": 8, "CONE": 9, "CONVEX_MESH": 10, "GAUSSIAN": 11}
UNLIMITED = 1e10
IDENTITY = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0]

# Newton's ShapeConfig defaults; gap is ModelBuilder.rigid_gap (builder.js:19-20).
SHAPE_DEFAULTS = dict(density=1000.0, mu=1.0, restitution=0.0, muTorsional=0.005, muRolling=0.0001, gap=0.1,
                      ke=2.5e3, kd=100.0, collisionGroup=1, world=-1, color=[0.7, 0.7, 0.7])
# Newton's JointDofConfig defaults (builder.js:22-23).
DOF_DEFAULTS = dict(limitLower=-UNLIMITED, limitUpper=UNLIMITED, targetPos=0.0, targetVel=0.0, targetKe=0.0,
                    targetKd=0.0, damping=0.0, armature=0.0, effortLimit=1e6, velocityLimit=1e6, friction=0.0)


# --- small rigid-transform math (double precision, like the JS host) -----------------------------
def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _add(a, b):
    return [a[0] + b[0], a[1] + b[1], a[2] + b[2]]


def _sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def _scale(a, s):
    return [a[0] * s, a[1] * s, a[2] * s]


def _norm(a):
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def qmul(a, b):
    return [a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1], a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
            a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3], a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2]]


def qconj(q):
    return [-q[0], -q[1], -q[2], q[3]]


def qrot(q, v):
    u = q[:3]
    t = _scale(_cross(u, v), 2.0)
    return _add(_add(list(v), _scale(t, q[3])), _cross(u, t))


def qaxis(axis, angle):
    n = _norm(axis)
    s = math.sin(angle / 2) / n
    return [axis[0] * s, axis[1] * s, axis[2] * s, math.cos(angle / 2)]


def tmul(a, b):
    return _add(a[:3], qrot(a[3:], b[:3])) + qmul(a[3:], b[3:])


def tinv(t):
    qi = qconj(t[3:])
    return _scale(qrot(qi, t[:3]), -1.0) + qi


def qmat(q):
    x, y, z, w = q
    return [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]


# --- convex hulls ---------------------------------------------------------------------------------
_HULL_CACHE: dict = {}


def _sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def _add(a, b):
    return [a[0] + b[0], a[1] + b[1], a[2] + b[2]]


def _scale(a, s):
    return [a[0] * s, a[1] * s, a[2] * s]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _hypot3(a):
    """Math.hypot of three numbers as the JS engines compute it: scaled by the largest magnitude, squares summed
    with Kahan compensation, then sqrt times the scale (so the hull planes round like the reference's)."""
    m = max(abs(a[0]), abs(a[1]), abs(a[2]))
    if m == 0.0:
        return 0.0
    total = comp = 0.0
    for v in a:
        x = abs(v) / m
        summand = x * x - comp
        pre = total + summand
        comp = (pre - total) - summand
        total = pre
    return math.sqrt(total) * m


class _Face:
    __slots__ = ("v", "n", "d", "alive", "out", "vis", "group")

    def __init__(self, P, a, b, c):
        n = _cross(_sub(P[b], P[a]), _sub(P[c], P[a]))
        n = _scale(n, 1.0 / _hypot3(n))
        self.v, self.n, self.d = (a, b, c), n, _dot(n, P[a])
        self.alive, self.out, self.vis, self.group = True, [], False, None


def convexHull(points):
    """Convex hull of a point cloud: {'verts': [[x,y,z]..], 'faces': [{'n','d','v'}..]}, outward polygon faces
    (coplanar triangles merged, vertices counter-clockwise about the normal, plane n.x = d).

    Follows the reference's incremental quickhull step for step (builder.js:61-173), in double precision like
    the JS host, so a SceneBuilder scene gets the reference's vertex order, face order and plane floats (which
    set feature ids and contact keys downstream): initial tetrahedron from extreme points; each point waits on
    the face it is farthest outside (beyond eps); the first live face with waiting points grows the hull through
    the faces its farthest point sees, breadth first over directed edges; connected coplanar triangles merge by a
    flood fill, each plane keeping its first triangle's (n, d); each face's loop is the 2D monotone-chain hull of
    its points; vertices are the used points in index order."""
    key = tuple(tuple(float(c) for c in p) for p in points)
    hit = _HULL_CACHE.get(key)
    if hit is not None:
        return hit
    seen = {}
    for p in key:
        if p not in seen:  # -0.0 and 0.0 are one key, as in the reference's string keys
            seen[p] = [p[0], p[1], p[2]]
    P = list(seen.values())
    if len(P) < 4:
        raise ValueError("a convex hull needs at least 4 distinct points")
    lo = [math.inf] * 3
    hi = [-math.inf] * 3
    for p in P:
        lo = [min(lo[k], p[k]) for k in range(3)]
        hi = [max(hi[k], p[k]) for k in range(3)]
    eps = 1e-7 * max(_hypot3(_sub(hi, lo)), 1e-12)
    i0 = 0
    for i in range(1, len(P)):
        if P[i][0] < P[i0][0]:
            i0 = i

    def far(fn):
        best, bd = -1, -1.0
        for i, p in enumerate(P):
            d = fn(p)
            if d > bd:
                bd, best = d, i
        return best, bd

    i1, _ = far(lambda p: _hypot3(_sub(p, P[i0])))
    e = _sub(P[i1], P[i0])
    ne = _hypot3(e)
    i2, d2 = far(lambda p: _hypot3(_cross(e, _sub(p, P[i0]))) / ne)
    if d2 <= eps:
        raise ValueError("convex hull points are collinear")
    n0 = _cross(e, _sub(P[i2], P[i0]))
    nn0 = _hypot3(n0)
    i3, d3 = far(lambda p: abs(_dot(n0, _sub(p, P[i0]))) / nn0)
    if d3 <= eps:
        raise ValueError("convex hull points are coplanar")
    center = _scale(_add(_add(P[i0], P[i1]), _add(P[i2], P[i3])), 0.25)
    K = 1048576
    edges = {}  # directed edge a * K + b -> face

    def make(a, b, c):
        f = _Face(P, a, b, c)
        for k in range(3):
            edges[f.v[k] * K + f.v[(k + 1) % 3]] = f
        return f

    def oriented(a, b, c):
        f = _Face(P, a, b, c)
        return (a, c, b) if _dot(f.n, center) - f.d > 0 else (a, b, c)

    faces = [make(*oriented(*t)) for t in ((i0, i1, i2), (i0, i1, i3), (i0, i2, i3), (i1, i2, i3))]

    def assign(pts, cands):
        for i in pts:
            best, bd = None, eps
            for f in cands:
                d = _dot(f.n, P[i]) - f.d
                if d > bd:
                    bd, best = d, f
            if best is not None:
                best.out.append(i)

    assign([i for i in range(len(P)) if i not in (i0, i1, i2, i3)], faces)
    for _ in range(4 * len(P) + 16):
        f0 = next((f for f in faces if f.alive and f.out), None)
        if f0 is None:
            break
        p, pd = f0.out[0], -math.inf
        for i in f0.out:
            d = _dot(f0.n, P[i]) - f0.d
            if d > pd:
                pd, p = d, i
        f0.vis = True
        visible, seen_f, horizon = [f0], {id(f0)}, []
        k = 0
        while k < len(visible):
            f = visible[k]
            k += 1
            for ei in range(3):
                a, b = f.v[ei], f.v[(ei + 1) % 3]
                nb = edges.get(b * K + a)
                if nb is None or id(nb) in seen_f:
                    if nb is not None and not nb.vis:
                        horizon.append((a, b))
                    continue
                if _dot(nb.n, P[p]) - nb.d > eps:
                    nb.vis = True
                    seen_f.add(id(nb))
                    visible.append(nb)
                else:
                    horizon.append((a, b))
        orphans = []
        for f in visible:
            f.alive = False
            for kk in range(3):
                ek = f.v[kk] * K + f.v[(kk + 1) % 3]
                if edges.get(ek) is f:
                    del edges[ek]
            orphans.extend(i for i in f.out if i != p)
        keep = [(a, b) for a, b in horizon if a * K + b not in edges]
        fresh = [make(a, b, p) for a, b in keep]
        assign(orphans, fresh)
        faces = [f for f in faces if f.alive] + fresh
    # Coplanar triangles merge into polygons by a flood fill across shared edges; a plane keeps its first
    # triangle's (n, d).
    planes = []
    for f in faces:
        if f.group is not None:
            continue
        pl = {"n": f.n, "d": f.d, "ids": {}}
        f.group = pl
        stack = [f]
        while stack:
            g = stack.pop()
            for v in g.v:
                pl["ids"].setdefault(v, None)  # insertion-ordered set
            for kk in range(3):
                nb = edges.get(g.v[(kk + 1) % 3] * K + g.v[kk])
                if (nb is not None and nb.group is None and _dot(nb.n, pl["n"]) > 0
                        and all(abs(_dot(pl["n"], P[v]) - pl["d"]) <= 10 * eps for v in nb.v)):
                    nb.group = pl
                    stack.append(nb)
        planes.append(pl)
    loops = []
    for pl in planes:
        ids = list(pl["ids"])
        acc = [0.0, 0.0, 0.0]
        for i in ids:
            acc = _add(acc, P[i])
        c = _scale(acc, 1.0 / len(ids))
        n = pl["n"]
        ref = [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0]
        u = _cross(n, ref)
        u = _scale(u, 1.0 / _hypot3(u))
        w = _cross(n, u)
        pts = sorted(((i, _dot(_sub(P[i], c), u), _dot(_sub(P[i], c), w)) for i in ids), key=lambda q: (q[1], q[2]))

        def turn(o, a, b):
            return (a[1] - o[1]) * (b[2] - o[2]) - (a[2] - o[2]) * (b[1] - o[1])

        def half(lst):
            h = []
            for q in lst:
                while len(h) >= 2 and turn(h[-2], h[-1], q) <= eps * eps:
                    h.pop()
                h.append(q)
            h.pop()
            return h

        ring = half(pts) + half(list(reversed(pts)))
        if len(ring) >= 3:
            loops.append({"n": n, "d": pl["d"], "ids": [q[0] for q in ring]})
    used = sorted({i for lp in loops for i in lp["ids"]})
    remap = {v: i for i, v in enumerate(used)}
    out = {"verts": [list(P[v]) for v in used],
           "faces": [{"n": list(lp["n"]), "d": lp["d"], "v": [remap[i] for i in lp["ids"]]} for lp in loops]}
    _HULL_CACHE[key] = out
    return out


def shapePoints(type_, s):
    """Hull sample points of the analytic shapes (builder.js:176-196)."""
    pts = []
    if type_ == GT["BOX"]:
        for x in (-1, 1):
            for y in (-1, 1):
                for z in (-1, 1):
                    pts.append([x * s[0], y * s[1], z * s[2]])
    elif type_ == GT["CYLINDER"]:
        for z in (-s[1], s[1]):
            for i in range(24):
                a = i * 2 * math.pi / 24
                pts.append([s[0] * math.cos(a), s[0] * math.sin(a), z])
    elif type_ == GT["CAPSULE"]:
        for z0, sg in ((s[1], 1), (-s[1], -1)):
            for e in range(4):
                el = e * math.pi / 6
                for k in range(12):
                    az = k * 2 * math.pi / 12
                    pts.append([s[0] * math.cos(el) * math.cos(az), s[0] * math.cos(el) * math.sin(az), z0 + sg * s[0] * math.sin(el)])
    elif type_ == GT["SPHERE"]:
        for e in range(7):
            el = -math.pi / 2 + e * math.pi / 6
            for k in range(12):
                az = k * 2 * math.pi / 12
                pts.append([s[0] * math.cos(el) * math.cos(az), s[0] * math.cos(el) * math.sin(az), s[0] * math.sin(el)])
    # Round to 1e-9 like the exporter, so hulls of equal shapes share one entry.
    return [[round(v * 1e9) / 1e9 + 0.0 for v in p] for p in pts]


def massProps(type_, s, hull):
    """Mass, COM and inertia about it of a unit-density solid (builder.js:199-238)."""
    def diag(m, a, b, c):
        return {"m": m, "c": [0.0, 0.0, 0.0], "I": [a, 0, 0, 0, b, 0, 0, 0, c]}
    if type_ == GT["BOX"]:
        m = 8 * s[0] * s[1] * s[2]
        return diag(m, m / 3 * (s[1] ** 2 + s[2] ** 2), m / 3 * (s[0] ** 2 + s[2] ** 2), m / 3 * (s[0] ** 2 + s[1] ** 2))
    if type_ == GT["SPHERE"]:
        m = 4 / 3 * math.pi * s[0] ** 3
        i = 0.4 * m * s[0] ** 2
        return diag(m, i, i, i)
    if type_ == GT["CYLINDER"]:
        m = math.pi * s[0] ** 2 * 2 * s[1]
        t = m * (3 * s[0] ** 2 + 4 * s[1] ** 2) / 12
        return diag(m, t, t, 0.5 * m * s[0] ** 2)
    if type_ == GT["CAPSULE"]:
        r, L = s[0], 2 * s[1]
        mc = math.pi * r * r * L
        ms = 4 / 3 * math.pi * r ** 3
        t = mc * (L * L / 12 + r * r / 4) + ms * (0.4 * r * r + L * L / 4 + 3 * L * r / 8)
        return diag(mc + ms, t, t, 0.5 * mc * r * r + 0.4 * ms * r * r)
    # Convex hull: tetrahedra from the vertex mean over each face fan.
    V = hull["verts"]
    o = _scale([sum(v[k] for v in V) for k in range(3)], 1 / len(V))
    m = 0.0
    c = [0.0, 0.0, 0.0]
    C = [0.0] * 9
    for f in hull["faces"]:
        for k in range(1, len(f["v"]) - 1):
            a = _sub(V[f["v"][0]], o)
            b = _sub(V[f["v"][k]], o)
            d = _sub(V[f["v"][k + 1]], o)
            vol = _dot(a, _cross(b, d)) / 6
            m += vol
            c = _add(c, _scale(_add(_add(a, b), d), vol / 4))
            for i in range(3):
                for j in range(3):
                    C[i * 3 + j] += vol / 20 * (2 * (a[i] * a[j] + b[i] * b[j] + d[i] * d[j]) + a[i] * b[j] + a[j] * b[i]
                                                + a[i] * d[j] + a[j] * d[i] + b[i] * d[j] + b[j] * d[i])
    c = _scale(c, 1 / m)
    tr = C[0] + C[4] + C[8]
    I = [(tr if i == j else 0.0) - C[i * 3 + j] for i in range(3) for j in range(3)]
    for i in range(3):
        for j in range(3):
            I[i * 3 + j] -= m * ((_dot(c, c) if i == j else 0.0) - c[i] * c[j])
    return {"m": m, "c": _add(c, o), "I": I}


class _Blob:
    """scene.bin writer: 4-byte float32/uint32 ranges {off, n}."""

    def __init__(self):
        self.chunks = []
        self.size = 0

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        off = self.size
        self.chunks.append(arr.tobytes())
        self.size += arr.nbytes
        return {"off": off, "n": int(arr.size)}

    def bytes(self):
        return b"".join(self.chunks)


class LoadedScene:
    """Scene json + binary blob, the object `Sim` consumes (sceneTypes.ts:91-99)."""

    def __init__(self, json_obj, bin_bytes, base=None):
        self.json = json_obj
        self.bin = bin_bytes
        self.base = base

    def f32(self, r):
        return np.frombuffer(self.bin, dtype=np.float32, count=r["n"], offset=r["off"])

    def u32(self, r):
        return np.frombuffer(self.bin, dtype=np.uint32, count=r["n"], offset=r["off"])

    def save(self, folder):
        os.makedirs(folder, exist_ok=True)
        with open(os.path.join(folder, "scene.json"), "w") as f:
            json.dump(self.json, f)
        with open(os.path.join(folder, "scene.bin"), "wb") as f:
            f.write(self.bin)


def loadScene(base):
    """Read a scene folder (scene.json + scene.bin), as sim.js:40-61 does (visual meshes are not read)."""
    with open(os.path.join(base, "scene.json")) as f:
        js = json.load(f)
    with open(os.path.join(base, "scene.bin"), "rb") as f:
        blob = f.read()
    return LoadedScene(js, blob, base)


class SceneBuilder:
    """Newton ModelBuilder analogue with the reference's API (builder.js:241-490)."""

    def __init__(self, gravity=(0.0, 0.0, -9.81), fps=60):
        self.gravity = list(gravity)
        self.fps = fps
        self.bodies = []
        self.shapes = []
        self.joints = []
        self.filters = []
        self.dofOf = []

    def addBody(self, xform=IDENTITY, mass=None, com=None, inertia=None, label=None, kinematic=False, gravity=True,
                twist=(0, 0, 0, 0, 0, 0)):
        self.bodies.append(dict(xform=list(map(float, xform)), mass=mass, com=com, inertia=inertia,
                                label=label if label is not None else f"body_{len(self.bodies)}", kinematic=kinematic,
                                gravity=gravity, twist=list(map(float, twist))))
        return len(self.bodies) - 1

    def setBodyInertia(self, body, mass, com, inertia):
        self.bodies[body].update(mass=mass, com=com, inertia=inertia)

    def addShape(self, body, type_, size, opts=None, hullPoints=None):
        o = dict(SHAPE_DEFAULTS)
        o.update(opts or {})
        hull = None if type_ == GT["PLANE"] else convexHull(hullPoints if hullPoints is not None else shapePoints(type_, size))
        self.shapes.append(dict(body=body, type=type_, size=list(size), xform=list(map(float, o.get("xform", IDENTITY))), hull=hull, o=o))
        return len(self.shapes) - 1

    def addShapeBox(self, body, hx=0.5, hy=0.5, hz=0.5, **o):
        return self.addShape(body, GT["BOX"], [hx, hy, hz], o)

    def addShapeSphere(self, body, radius=0.5, **o):
        return self.addShape(body, GT["SPHERE"], [radius, 0, 0], o)

    def addShapeCapsule(self, body, radius=0.05, halfHeight=0.5, **o):
        return self.addShape(body, GT["CAPSULE"], [radius, halfHeight, 0], o)

    def addShapeCylinder(self, body, radius=0.05, halfHeight=0.5, **o):
        return self.addShape(body, GT["CYLINDER"], [radius, halfHeight, 0], o)

    def addShapeConvexHull(self, body, vertices, **o):
        pts = [list(v) for v in vertices] if np.ndim(vertices[0]) else [list(vertices[3 * i:3 * i + 3]) for i in range(len(vertices) // 3)]
        return self.addShape(body, GT["CONVEX_MESH"], [1, 1, 1], o, pts)

    def addGroundPlane(self, height=0.0, **o):
        opts = dict(xform=[0, 0, height, 0, 0, 0, 1])
        opts.update(o)
        return self.addShape(-1, GT["PLANE"], [0, 0, 0], opts)

    def addCollisionFilter(self, shapeA, shapeB):
        self.filters.append([shapeA, shapeB])

    def addJoint(self, type_, parent, child, parentXform=IDENTITY, childXform=IDENTITY, axes=(), linear=0, q=None,
                 label=None, collisionFilterParent=True, dofs=None, **dof):
        base = dict(DOF_DEFAULTS)
        base.update(dof)
        self.joints.append(dict(type=type_, parent=parent, child=child, Xp=list(map(float, parentXform)), Xc=list(map(float, childXform)),
                                axes=[list(map(float, a)) for a in axes], linear=linear, q=q, label=label, filter=collisionFilterParent,
                                dofs=dofs if dofs is not None else [dict(base) for _ in axes], mimic=None))
        return len(self.joints) - 1

    def addJointRevolute(self, parent, child, axis=(1, 0, 0), **o):
        return self.addJoint(JT["REVOLUTE"], parent, child, axes=[axis], **o)

    def addJointPrismatic(self, parent, child, axis=(1, 0, 0), **o):
        return self.addJoint(JT["PRISMATIC"], parent, child, axes=[axis], linear=1, **o)

    def addJointBall(self, parent, child, **o):
        return self.addJoint(JT["BALL"], parent, child, axes=[[1, 0, 0], [0, 1, 0], [0, 0, 1]], **o)

    def addJointFixed(self, parent, child, **o):
        return self.addJoint(JT["FIXED"], parent, child, **o)

    def addJointFree(self, child, **o):
        return self.addJoint(JT["FREE"], -1, child, axes=[[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], linear=3, **o)

    def setJointDofConfig(self, joint, axisIndex, cfg):
        self.joints[joint]["dofs"][axisIndex].update(cfg)

    def jointDof(self, joint):
        return self.dofOf[joint]

    def finalize(self, solver=None):
        """Scene json + blob in the reference's format (builder.js:320-490)."""
        NB = len(self.bodies)
        parentJoint = [-1] * NB
        loop = set()
        for i, j in enumerate(self.joints):
            if j["child"] < 0 or j["child"] >= NB:
                raise ValueError(f"joint {i}: bad child body {j['child']}")
            if parentJoint[j["child"]] >= 0:
                loop.add(i)
            else:
                parentJoint[j["child"]] = i
        order, art, implicit, placed = [], [], [], set()

        def rootOf(b):
            r, n = b, 0
            while parentJoint[r] >= 0 and self.joints[parentJoint[r]]["parent"] >= 0 and n < NB:
                r = self.joints[parentJoint[r]]["parent"]
                n += 1
            return r

        rootof = [rootOf(b) for b in range(NB)]
        roots = sorted(set(rootof))
        jroot = [rootOf(j["child"]) for j in self.joints]
        byroot = {}  # each root's tree joints in joint order
        for i in range(len(self.joints)):
            if i not in loop:
                byroot.setdefault(jroot[i], []).append(i)
        for a, r in enumerate(roots):
            if parentJoint[r] < 0:
                implicit.append(dict(type=JT["FREE"], parent=-1, child=r, Xp=IDENTITY, Xc=IDENTITY,
                                     axes=[[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], linear=3,
                                     label=f"{self.bodies[r]['label']}_free", dofs=[dict(DOF_DEFAULTS) for _ in range(6)], mimic=None))
                order.append(-len(implicit))
                art.append(a)
            for i in byroot.get(r, ()):
                j = self.joints[i]
                order.append(i)
                art.append(a)
                placed.add(j["child"])
            placed.add(r)
        for i in sorted(loop):
            order.append(i)
            art.append(-1)

        def jointAt(k):
            return implicit[-k - 1] if k < 0 else self.joints[k]

        J = dict(type=[], parent=[], child=[], X_p=[], X_c=[], axis=[], dof_dim=[], q_start=[0], qd_start=[0], articulation=art,
                 q=[], qd=[], armature=[], ke=[], kd=[], effort=[], lower=[], upper=[], vel_limit=[], friction=[], passive_damping=[],
                 target_q=[], target_qd=[], mimic_joint=[], mimic_coeffs=[], label=[])
        self.dofOf = [0] * len(self.joints)
        for idx, k in enumerate(order):
            j = jointAt(k)
            if k >= 0:
                self.dofOf[k] = J["qd_start"][idx]
            ndof = len(j["axes"])
            scalar = j["type"] in (JT["REVOLUTE"], JT["PRISMATIC"], JT["D6"])
            if j["type"] == JT["FREE"]:
                q = list(self.bodies[j["child"]]["xform"])
            elif j["type"] == JT["BALL"]:
                q = list(j["q"]) if j["q"] is not None else [0, 0, 0, 1]
            elif scalar:
                q0 = j["q"] if j["q"] is not None else 0
                q = (list(q0) if isinstance(q0, (list, tuple)) else [q0]) + [0] * ndof
                q = q[:ndof]
            else:
                q = []
            J["type"].append(j["type"]); J["parent"].append(j["parent"]); J["child"].append(j["child"])
            J["X_p"].append(list(j["Xp"])); J["X_c"].append(list(j["Xc"]))
            J["dof_dim"].append([j["linear"], ndof - j["linear"]])
            J["q"] += [float(v) for v in q]
            J["target_q"] += [d["targetPos"] for d in j["dofs"]] if scalar else [0.0] * len(q)
            J["q_start"].append(len(J["q"]))
            J["qd"] += list(self.bodies[j["child"]]["twist"]) if j["type"] == JT["FREE"] else [0.0] * ndof
            for a, ax in enumerate(j["axes"]):
                d = j["dofs"][a]
                J["axis"].append(list(ax)); J["armature"].append(d["armature"]); J["ke"].append(d["targetKe"]); J["kd"].append(d["targetKd"])
                J["effort"].append(d["effortLimit"]); J["lower"].append(d["limitLower"]); J["upper"].append(d["limitUpper"])
                J["vel_limit"].append(d["velocityLimit"]); J["friction"].append(d["friction"]); J["passive_damping"].append(d["damping"])
                J["target_qd"].append(d["targetVel"])
            J["qd_start"].append(len(J["axis"]))
            J["mimic_joint"].append(-1)
            J["mimic_coeffs"].append([0, 1])
            J["label"].append(j["label"] if j["label"] is not None else f"joint_{idx}")
        J["count"] = len(order)

        # Forward kinematics for the initial body poses.
        pose = [list(b["xform"]) for b in self.bodies]
        for idx, k in enumerate(order):
            if art[idx] < 0:
                continue
            j = jointAt(k)
            if j["type"] == JT["FREE"]:
                continue
            qs = J["q_start"][idx]
            T = list(IDENTITY)
            if j["type"] == JT["REVOLUTE"]:
                T = [0.0, 0.0, 0.0] + qaxis(j["axes"][0], J["q"][qs])
            elif j["type"] == JT["PRISMATIC"]:
                T = _scale(j["axes"][0], J["q"][qs] / _norm(j["axes"][0])) + [0.0, 0.0, 0.0, 1.0]
            elif j["type"] == JT["BALL"]:
                T = [0.0, 0.0, 0.0] + J["q"][qs:qs + 4]
            elif j["type"] == JT["D6"]:
                p, r = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0]
                for a in range(j["linear"]):
                    p = _add(p, _scale(j["axes"][a], J["q"][qs + a]))
                for a in range(j["linear"], len(j["axes"])):
                    ax = qrot(r, _scale(j["axes"][a], 1 / _norm(j["axes"][a])))
                    r = qmul(qaxis(ax, J["q"][qs + a]), r)
                T = p + r
            Xwp = pose[j["parent"]] if j["parent"] >= 0 else IDENTITY
            pose[j["child"]] = tmul(tmul(tmul(Xwp, j["Xp"]), T), tinv(j["Xc"]))

        bodies = dict(count=NB, mass=[], inertia=[], com=[], q=pose, twist=[], label=[], kinematic=[], no_gravity=[])
        shapes_of = {}  # each body's shapes in shape order
        for s in self.shapes:
            shapes_of.setdefault(s["body"], []).append(s)
        for i, b in enumerate(self.bodies):
            m, c, parts = 0.0, [0.0, 0.0, 0.0], []
            for s in shapes_of.get(i, ()):
                if s["type"] == GT["PLANE"] or not ((s["o"].get("mass") or 0) > 0 or s["o"]["density"] > 0):
                    continue
                p = massProps(s["type"], s["size"], s["hull"])
                rho = s["o"]["mass"] / p["m"] if (s["o"].get("mass") or 0) > 0 else s["o"]["density"]
                pm = p["m"] * rho
                pc = _add(s["xform"][:3], qrot(s["xform"][3:], p["c"]))
                parts.append(dict(m=pm, c=pc, I=[v * rho for v in p["I"]], R=qmat(s["xform"][3:])))
                m += pm
                c = _add(c, _scale(pc, pm))
            c = _scale(c, 1 / m) if m > 0 else [0.0, 0.0, 0.0]
            I = [0.0] * 9
            for p in parts:
                d = _sub(p["c"], c)
                for r in range(3):
                    for s_ in range(3):
                        v = 0.0
                        for a in range(3):
                            for e in range(3):
                                v += p["R"][r * 3 + a] * p["I"][a * 3 + e] * p["R"][s_ * 3 + e]
                        I[r * 3 + s_] += v + p["m"] * ((_dot(d, d) if r == s_ else 0.0) - d[r] * d[s_])
            bodies["mass"].append(b["mass"] if b["mass"] is not None else m)
            bodies["com"].append(list(b["com"]) if b["com"] is not None else c)
            if b["inertia"] is not None:
                bodies["inertia"].append(list(b["inertia"]))
            else:
                bodies["inertia"].append([v * b["mass"] / m for v in I] if (b["mass"] is not None and m > 0) else I)
            bodies["twist"].append(list(b["twist"])); bodies["label"].append(b["label"])
            bodies["kinematic"].append(bool(b["kinematic"])); bodies["no_gravity"].append(not b["gravity"])

        blob = _Blob()
        shapes, index = [], {}
        for sid, s in enumerate(self.shapes):
            o = s["o"]
            entry = dict(sid=sid, body=s["body"], xf=list(s["xform"]), type=s["type"], color=o["color"], mu=o["mu"], gap=o["gap"],
                         restitution=o["restitution"], mu_torsional=o["muTorsional"], mu_rolling=o["muRolling"], group=o["collisionGroup"],
                         condim=3, ke=o["ke"], kd=o["kd"], torsion_radius=0, world=o["world"])
            if s["type"] == GT["PLANE"]:
                entry["plane"] = True
            else:
                fv, fstart, fn = [], [], []
                for f in s["hull"]["faces"]:
                    fstart.append(len(fv)); fv += f["v"]; fn += list(f["n"]) + [f["d"]]
                entry["hull"] = dict(verts=blob.put(np.asarray(s["hull"]["verts"], np.float32).ravel()),
                                     face_planes=blob.put(np.asarray(fn, np.float32)),
                                     face_start=blob.put(np.asarray(fstart + [len(fv)], np.uint32)),
                                     face_verts=blob.put(np.asarray(fv, np.uint32)), nv=len(s["hull"]["verts"]), nf=len(s["hull"]["faces"]))
                if s["type"] == GT["SPHERE"]:
                    entry["prim"] = dict(kind=1, r=s["size"][0], h=0)
                if s["type"] == GT["CAPSULE"]:
                    entry["prim"] = dict(kind=2, r=s["size"][0], h=s["size"][1])
            index[sid] = len(shapes)
            shapes.append(entry)
        # Newton's collision_filter_parent: a joint's two bodies do not collide (builder.js:470-477).
        pairs = {(min(a, b), max(a, b)) for a, b in self.filters}
        by_body = {}
        for si, s in enumerate(self.shapes):
            by_body.setdefault(s["body"], []).append(si)
        for j in self.joints:
            if not j["filter"] or j["parent"] < 0:
                continue
            for a in by_body.get(j["parent"], []):
                for b in by_body.get(j["child"], []):
                    pairs.add((min(a, b), max(a, b)))
        targets = blob.put(np.zeros(0, np.float32))
        s = dict(substeps=4, iterations=12, enable_joint_limits=True)
        s.update(solver or {})
        js = dict(gravity=self.gravity, fps=self.fps, substeps=s["substeps"], iterations=s["iterations"], bodies=bodies, joints=J,
                  shapes=shapes, visuals=[], filters=[[index[a], index[b]] for a, b in sorted(pairs)], mimics=[], targets=targets,
                  target_frames=0, driven=0, integral_count=0, demo="built", control=None, solver=s, info={},
                  joint_type_names=dict(JT), geo_type_names=dict(GT))
        return LoadedScene(js, blob.bytes())
