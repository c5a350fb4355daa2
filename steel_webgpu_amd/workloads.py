"""Synthetic workloads: the benchmark scenes, built with SceneBuilder from a seeded PCG32 stream.

Scenes (fps / substeps / steps as in the benchmark definitions):
  boxes64        ground plane + 64 boxes, half extents U[0.05,0.2]^3, density 1000, mu 0.6; 240 / 4 / 500
  arm_hulls      32-link fixed-base revolute capsule chain among 256 convex hulls in a walled 2x2 m bin; 240 / 8 / 1000
  quadruped(i)   floating-base 12-joint quadruped + K~Geom(0.1)<=64 boxes, world i of a batch; 200 / 4 / 1000
  large_world    one world of 100k bodies (boxes and 12-point sphere hulls, sizes U[0.03,0.1]) on a jittered 3D lattice
                 over a 20x20 m floor, plus 256 quadrupeds; 240 / 4 / 300
Gravity is -z and the ground is z = 0, so the vertical range of each spawn box is its z range.
"""
from __future__ import annotations

import math

import numpy as np

from .scene import GT, SceneBuilder

GT_BOX = GT["BOX"]


class PCG32:
    """PCG-XSH-RR 64/32 (O'Neill 2014): seeded, stream-selectable, platform independent."""

    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed, stream=0):
        self.inc = ((stream << 1) | 1) & self.MASK
        self.state = 0
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * (self.next_u32() / 4294967296.0)

    def geometric(self, p):
        """Failures before the first success, P(k) = (1-p)^k p."""
        u = 1.0 - self.uniform()
        return int(math.floor(math.log(u) / math.log(1.0 - p)))

    def unit_quat(self):
        """Uniform random rotation (Shoemake, Graphics Gems III), as x, y, z, w."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1 - u1), math.sqrt(u1)
        return [a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)]

    def unit_vec(self):
        z = self.uniform(-1, 1)
        t = self.uniform(0, 2 * math.pi)
        r = math.sqrt(max(0.0, 1 - z * z))
        return [r * math.cos(t), r * math.sin(t), z]


SEED = 0x57EE1000


def _place(rng, lo, hi, radius, placed, tries=1000, axis_clear=0.0):
    """Uniform position in the box [lo, hi], re-drawn while its bounding sphere overlaps a placed one, reaches
    below the ground (z = 0), or comes within axis_clear of the vertical axis x = y = 0."""
    for _ in range(tries):
        p = [rng.uniform(lo[k], hi[k]) for k in range(3)]
        if axis_clear and p[0] ** 2 + p[1] ** 2 < (axis_clear + radius) ** 2:
            continue
        if p[2] >= radius and all((p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 + (p[2] - q[2]) ** 2 >= (radius + r) ** 2 for q, r in placed):
            placed.append((p, radius))
            return p
    raise RuntimeError("no free spawn position")


def _qmat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _obb_overlap(c1, R1, h1, c2, R2, h2):
    """Exact separating-axis test of two oriented boxes (15 axes)."""
    d = c2 - c1
    axes = [R1[:, i] for i in range(3)] + [R2[:, i] for i in range(3)]
    axes += [np.cross(R1[:, i], R2[:, j]) for i in range(3) for j in range(3)]
    for a in axes:
        n = np.linalg.norm(a)
        if n < 1e-9:
            continue
        a = a / n
        r1 = sum(h1[i] * abs(np.dot(R1[:, i], a)) for i in range(3))
        r2 = sum(h2[i] * abs(np.dot(R2[:, i], a)) for i in range(3))
        if abs(np.dot(d, a)) > r1 + r2:
            return False
    return True


def _place_box(rng, lo, hi, h, q, boxes, tries=20000):
    """Position in [lo, hi] for a box (half extents h, rotation q), re-drawn while it overlaps a placed
    box (exact oriented-box test) or reaches below the ground."""
    R = _qmat(q)
    hz = sum(h[i] * abs(R[2, i]) for i in range(3))
    hv = np.asarray(h, float)
    rad = float(np.linalg.norm(hv))
    for _ in range(tries):
        c = np.array([rng.uniform(lo[k], hi[k]) for k in range(3)])
        if c[2] - hz < 0.0:
            continue
        ok = True
        for c2, R2, h2, r2 in boxes:
            if np.sum((c - c2) ** 2) < (rad + r2) ** 2 and _obb_overlap(c, R, hv, c2, R2, h2):
                ok = False
                break
        if ok:
            boxes.append((c, R, hv, rad))
            return [float(v) for v in c]
    raise RuntimeError("no free spawn position")


def _add_box(b, rng, placed, lo, hi, mu):
    h = [rng.uniform(0.05, 0.2) for _ in range(3)]
    q = rng.unit_quat()
    p = _place_box(rng, lo, hi, h, q, placed)
    body = b.addBody(xform=p + q)
    b.addShapeBox(body, hx=h[0], hy=h[1], hz=h[2], mu=mu, density=1000.0)
    return body


def boxes64(seed=SEED + 0, n=64, fps=240, substeps=4, iterations=12):
    """Ground plane + n random boxes (the plain single-world scene)."""
    rng = PCG32(seed, 0)
    b = SceneBuilder(fps=fps)
    b.addGroundPlane(mu=0.6)
    placed = []
    for _ in range(n):
        _add_box(b, rng, placed, (-1, -1, 0.2), (1, 1, 2.2), 0.6)
    return b.finalize(solver=dict(substeps=substeps, iterations=iterations))


def arm_hulls(seed=SEED + 1, nlinks=32, nhulls=256, fps=240, substeps=8, iterations=12):
    """32-link fixed-base revolute capsule chain among convex hulls in a walled 2x2 m bin."""
    rng = PCG32(seed, 0)
    b = SceneBuilder(fps=fps)
    b.addGroundPlane(mu=0.6)
    for sx, sy, hx, hy in ((1.05, 0, 0.05, 1.1), (-1.05, 0, 0.05, 1.1), (0, 1.05, 1.1, 0.05), (0, -1.05, 1.1, 0.05)):
        b.addShapeBox(-1, hx=hx, hy=hy, hz=0.5, mu=0.6, xform=[sx, sy, 0.5, 0, 0, 0, 1])
    parent, prev_len, z = -1, 0.0, 0.0
    for i in range(nlinks):
        L = rng.uniform(0.1, 0.3)
        body = b.addBody(xform=[0, 0, z, 0, 0, 0, 1])
        b.addShapeCapsule(body, radius=0.03, halfHeight=L / 2, mu=0.6, xform=[0, 0, L / 2, 0, 0, 0, 1])
        axis = rng.unit_vec()
        b.addJointRevolute(parent, body, axis=axis, parentXform=[0, 0, prev_len if parent >= 0 else 0.0, 0, 0, 0, 1],
                           limitLower=-math.pi / 2, limitUpper=math.pi / 2)
        parent, prev_len, z = body, L, z + L
    placed = []
    for _ in range(nhulls):
        r = rng.uniform(0.04, 0.12)
        npts = min(8 + int(rng.uniform(0, 9)), 16)
        pts = []
        for _ in range(npts):
            v = rng.unit_vec()
            pts.append([r * v[0], r * v[1], r * v[2]])
        # clear of the standing arm (capsules of radius 0.03 on the vertical axis) by 2 cm
        p = _place(rng, (-0.9, -0.9, 0.15), (0.9, 0.9, 2.5), r, placed, axis_clear=0.05)
        body = b.addBody(xform=p + rng.unit_quat())
        b.addShapeConvexHull(body, vertices=pts, mu=0.6, density=1000.0)
    return b.finalize(solver=dict(substeps=substeps, iterations=iterations))


def quadruped_world(index, seed=SEED + 2, fps=200, substeps=4, iterations=12, kmax=64, drive=None, solver=None):
    """One world of the batch: floating-base quadruped (box base, 4 legs x 3 revolute capsule links) + K boxes.
    `drive` (dict of targetKe, targetKd, effortLimit, ...) gives every leg joint a PD drive holding its initial
    angle, `solver` adds solver options; the batch workload has neither."""
    rng = PCG32(seed, index)
    b = SceneBuilder(fps=fps)
    b.addGroundPlane(mu=0.6)
    base_z = 0.55
    _add_quadruped(b, 0.0, 0.0, rng, base_z, drive)
    K = min(rng.geometric(0.1), kmax)
    # keep-out box around the quadruped
    placed = [(np.array([0.0, 0.0, base_z - 0.2]), np.eye(3), np.array([0.45, 0.3, 0.45]), 0.71)]
    for _ in range(K):
        _add_box(b, rng, placed, (-1.5, -1.5, 0.2), (1.5, 1.5, 2.2), 0.6)
    return b.finalize(solver=dict(substeps=substeps, iterations=iterations, **(solver or {})))


# --- serving requests ----------------------------------------------------------------------------------------
_HULL_TEMPLATES = None


def hull_templates(n=64, seed=SEED + 3):
    """A fixed pool of convex hulls of 8-16 points on spheres of radius U[0.04, 0.12] (request hulls pick one)."""
    global _HULL_TEMPLATES
    if _HULL_TEMPLATES is None:
        rng = PCG32(seed, 0)
        out = []
        for _ in range(n):
            r = rng.uniform(0.04, 0.12)
            npts = min(8 + int(rng.uniform(0, 9)), 16)
            out.append([[r * c for c in rng.unit_vec()] for _ in range(npts)])
        _HULL_TEMPLATES = out
    return _HULL_TEMPLATES


def _add_quadruped(b, x0, y0, rng, base_z=0.55, drive=None):
    base = b.addBody(xform=[x0, y0, base_z, 0, 0, 0, 1])
    b.addShapeBox(base, hx=0.2, hy=0.05, hz=0.1, mu=0.6, density=1000.0)
    axes = ([1, 0, 0], [0, 1, 0], [0, 1, 0])
    for sx in (1, -1):
        for sy in (1, -1):
            parent = base
            pxf = [sx * 0.2, sy * 0.05, -0.1, 0, 0, 0, 1]
            for k in range(3):
                body = b.addBody(xform=[0, 0, 0, 0, 0, 0, 1])
                b.addShapeCapsule(body, radius=0.02, halfHeight=0.075, mu=0.6, xform=[0, 0, -0.075, 0, 0, 0, 1])
                q = rng.uniform(-0.3, 0.3)
                b.addJointRevolute(parent, body, axis=axes[k], parentXform=pxf, q=q, **(dict(drive, targetPos=q) if drive else {}))
                parent, pxf = body, [0, 0, -0.15, 0, 0, 0, 1]


def serving_request(index, seed=SEED + 3, fps=240, substeps=4, iterations=12, nmin=10, nmax=2000, hmin=50, hmax=1000):
    """One serving request: (scene, horizon).  N ~ log-uniform [nmin, nmax] bodies, 30% hulls / 70% boxes,
    plus 0-4 quadrupeds; horizon ~ U[hmin, hmax] frames.  Bodies start on a jittered 3D lattice (0.75 m cells,
    4 layers from z = 0.4 m) so no two overlap; quadrupeds stand in a row beside the lattice."""
    rng = PCG32(seed, 1000 + index)
    n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
    nq = min(int(rng.uniform(0, 5)), 4)
    horizon = int(rng.uniform(hmin, hmax + 1))
    b = SceneBuilder(fps=fps)
    b.addGroundPlane(mu=0.6)
    layers = 4 if n >= 16 else 1
    side = max(1, int(math.ceil(math.sqrt(n / layers))))
    cell = 0.75
    temps = hull_templates()
    for i in range(n):
        lz, rem = divmod(i, side * side)
        ly, lx = divmod(rem, side)
        p = [(lx - 0.5 * (side - 1)) * cell + rng.uniform(-0.02, 0.02), (ly - 0.5 * (side - 1)) * cell + rng.uniform(-0.02, 0.02),
             0.4 + lz * cell + rng.uniform(0.0, 0.02)]
        q = rng.unit_quat()
        body = b.addBody(xform=p + q)
        if rng.uniform() < 0.3:
            b.addShapeConvexHull(body, vertices=temps[int(rng.uniform(0, len(temps))) % len(temps)], mu=0.6, density=1000.0)
        else:
            h = [rng.uniform(0.05, 0.2) for _ in range(3)]
            b.addShapeBox(body, hx=h[0], hy=h[1], hz=h[2], mu=0.6, density=1000.0)
    x_row = -0.5 * side * cell - 1.5
    for k in range(nq):
        _add_quadruped(b, x_row, (k - 1.5) * 1.2, rng)
    return b.finalize(solver=dict(substeps=substeps, iterations=iterations)), horizon


def with_target_path(scene, path, integral=0, held=False):
    """Attach a recorded joint-target path to a scene: `path` is frames x driven (60 Hz samples of the targets of
    the first `driven` dofs, the scene-folder fields targets / target_frames / driven, sim.js:210, 372, 633),
    `integral` dofs get the path's integral term, `held` holds each frame's sample."""
    path = np.ascontiguousarray(path, np.float32)
    off = len(scene.bin)
    scene.bin = bytes(scene.bin) + path.tobytes()
    J = scene.json
    J["targets"] = {"off": off, "n": int(path.size)}
    J["target_frames"], J["driven"] = path.shape
    J["integral_count"], J["targets_held"] = integral, held
    return scene


def gait_path(frames=120, dofs=18, free=6, amp=0.3, hz=1.5, seed=0):
    """A 60 Hz sinusoidal gait for a floating-base quadruped: the first `free` dofs (the base) stay 0, each leg
    dof swings with its own phase."""
    rng = np.random.default_rng(seed)
    t = np.arange(frames) / 60.0
    ph = rng.uniform(0, 2 * np.pi, dofs)
    path = amp * np.sin(2 * np.pi * hz * t[:, None] + ph[None, :])
    path[:, :free] = 0.0
    return path


# --- the large single world -----------------------------------------------------------------------------------
LARGE_PITCH = 0.36   # lattice cell (m): two bodies of bounding radius <= 0.1 sqrt(3) never overlap with 5 mm jitter
LARGE_GAP = 0.02     # contact gap (m) of the lattice bodies: the shipped scenes' range (0-0.02), see large_world


def _hull_job(pts):
    from . import scene as _scene
    return _scene.convexHull(pts)


def large_world(seed=SEED + 4, nbodies=100000, nquads=256, extent=20.0, gap=LARGE_GAP, fps=240, substeps=4, iterations=12,
                workers=1):
    """One large world: `nbodies` bodies, alternately boxes (half extents U[0.03, 0.1] per axis) and convex hulls of 12
    points on a sphere of radius U[0.03, 0.1], density 1000, mu 0.6, random orientations, at rest on a jittered 3D
    lattice (LARGE_PITCH cells, +-5 mm) over an extent x extent m floor, filled layer by layer from z = 0.2 m; and
    `nquads` floating-base quadrupeds standing on a square grid over the same floor, whose lattice cells below 1.1 m
    are left empty.  The lattice bodies use a `gap` contact gap (the builder's default 0.1 m, against bodies of 3-10 cm,
    would make every lattice neighbour a speculative contact pair); the quadrupeds keep the default.  `workers` > 1
    computes the shapes' convex hulls in that many processes first (the same hulls: SceneBuilder finds them cached)."""
    rng = PCG32(seed, 0)
    b = SceneBuilder(fps=fps)
    b.addGroundPlane(mu=0.6)
    qs = max(1, int(math.ceil(math.sqrt(nquads)))) if nquads else 1
    qpitch = extent / qs
    quads = []
    for k in range(nquads):
        qy, qx = divmod(k, qs)
        x0, y0 = (qx + 0.5) * qpitch - 0.5 * extent, (qy + 0.5) * qpitch - 0.5 * extent
        _add_quadruped(b, x0, y0, rng)
        quads.append((x0, y0))
    side = max(1, int(extent / LARGE_PITCH))
    x_cells = [(i + 0.5) * LARGE_PITCH - 0.5 * side * LARGE_PITCH for i in range(side)]

    def blocked(x, y, z):
        if z >= 1.1:
            return False
        cx, cy = int(math.floor((x + 0.5 * extent) / qpitch)), int(math.floor((y + 0.5 * extent) / qpitch))
        for ix in (cx - 1, cx, cx + 1):
            for iy in (cy - 1, cy, cy + 1):
                k = iy * qs + ix
                if 0 <= ix < qs and 0 <= iy < qs and k < nquads:
                    qx, qy = quads[k]
                    if abs(x - qx) < 0.45 + 0.5 * LARGE_PITCH and abs(y - qy) < 0.3 + 0.5 * LARGE_PITCH:
                        return True
        return False

    bodies = []  # (pose, box half extents or None, hull points or None), drawn in lattice order
    layer = 0
    while len(bodies) < nbodies:
        z = 0.2 + layer * LARGE_PITCH
        for y in x_cells:
            for x in x_cells:
                if len(bodies) >= nbodies:
                    break
                if blocked(x, y, z):
                    continue
                p = [x + rng.uniform(-0.005, 0.005), y + rng.uniform(-0.005, 0.005), z + rng.uniform(-0.005, 0.005)]
                pose = p + rng.unit_quat()
                if len(bodies) % 2 == 0:
                    bodies.append((pose, [rng.uniform(0.03, 0.1) for _ in range(3)], None))
                else:
                    r = rng.uniform(0.03, 0.1)
                    bodies.append((pose, None, [[r * c for c in rng.unit_vec()] for _ in range(12)]))
        layer += 1
    if workers > 1:
        from . import scene as _scene
        jobs = [_scene.shapePoints(GT_BOX, h) if h is not None else pts for _, h, pts in bodies]
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            hulls = pool.map(_hull_job, jobs, chunksize=256)
        for pts, hull in zip(jobs, hulls):
            _scene._HULL_CACHE[tuple(tuple(float(c) for c in p) for p in pts)] = hull
    for pose, h, pts in bodies:
        body = b.addBody(xform=pose)
        if h is not None:
            b.addShapeBox(body, hx=h[0], hy=h[1], hz=h[2], mu=0.6, density=1000.0, gap=gap)
        else:
            b.addShapeConvexHull(body, vertices=pts, mu=0.6, density=1000.0, gap=gap)
    return b.finalize(solver=dict(substeps=substeps, iterations=iterations))


def large_world_small(seed=SEED + 4, nbodies=10000, nquads=25):
    """The large world's construction at a tenth of its bodies (same lattice density: the floor scaled by
    sqrt(1/10)), for parity tests the oracle finishes in seconds."""
    return large_world(seed=seed, nbodies=nbodies, nquads=nquads, extent=20.0 * math.sqrt(nbodies / 100000.0))


def offset_stack(nfar=0, seed=SEED + 5, fps=240, substeps=4, iterations=12):
    """Two boxes stacked on the ground, preceded by `nfar` free bodies without shapes (1 kg, falling far from the
    stack).  The stack's bodies are numbered nfar and nfar + 1 while its shapes keep the numbers 1 and 2, so with
    nfar >= 65535 its contact keys, units, colors and anchors need body ids past 16 bits, while everything keyed by
    shape (feature hashes) is unchanged: its trajectory is the same for every nfar."""
    rng = PCG32(seed, 0)
    b = SceneBuilder(fps=fps)
    b.addGroundPlane(mu=0.6)
    for i in range(nfar):
        b.addBody(xform=[100.0 + (i % 300), float(i // 300), 50.0, 0, 0, 0, 1], mass=1.0, com=[0.0, 0.0, 0.0],
                  inertia=[0.01, 0, 0, 0, 0.01, 0, 0, 0, 0.01])
    h = [[rng.uniform(0.05, 0.2) for _ in range(3)] for _ in range(2)]
    q = rng.unit_quat()
    lower = b.addBody(xform=[0.0, 0.0, h[0][2] + 0.01, 0, 0, 0.0998, 0.995])
    b.addShapeBox(lower, hx=h[0][0], hy=h[0][1], hz=h[0][2], mu=0.6, density=1000.0)
    upper = b.addBody(xform=[0.02, -0.01, 2 * h[0][2] + h[1][2] + 0.05] + q)
    b.addShapeBox(upper, hx=h[1][0], hy=h[1][1], hz=h[1][2], mu=0.6, density=1000.0)
    return b.finalize(solver=dict(substeps=substeps, iterations=iterations))


def cell_mix(seed=SEED + 6, n=4600, far=False, fps=240, substeps=4, iterations=12):
    """A world past the cell broadphase's shape count with shapes it bins and shapes it does not: `n` free boxes of
    half extents U[0.03, 0.15] at random poses in an 8 x 8 x 6 m region (overlaps allowed), six large free boxes and
    three large static boxes (tested against every shape), and the ground plane.  With `far`, two overlapping pairs
    of boxes sit 5e6 m out on either side of x, so the cell coordinates of everything else are clamped."""
    rng = PCG32(seed, 0)
    b = SceneBuilder(fps=fps)
    b.addGroundPlane(mu=0.6)
    for k in range(3):
        b.addShapeBox(-1, hx=3.0, hy=0.5 + k, hz=0.4, mu=0.6, xform=[-4.0 + 4.0 * k, 2.0, 0.4, 0, 0, 0, 1])
    for i in range(n + 6):
        big = i >= n
        hs = [rng.uniform(0.5, 0.8) if big else rng.uniform(0.03, 0.15) for _ in range(3)]
        p = [rng.uniform(-4.0, 4.0), rng.uniform(-4.0, 4.0), rng.uniform(0.2, 6.0)]
        body = b.addBody(xform=p + rng.unit_quat())
        b.addShapeBox(body, hx=hs[0], hy=hs[1], hz=hs[2], mu=0.6, density=500.0)
    if far:
        for x in (5e6, 5e6 + 0.15, -5e6, -5e6 - 0.15):
            body = b.addBody(xform=[x, 0.0, 5.0, 0, 0, 0, 1])
            b.addShapeBox(body, hx=0.1, hy=0.1, hz=0.1, mu=0.6, density=500.0)
    return b.finalize(solver=dict(substeps=substeps, iterations=iterations))
