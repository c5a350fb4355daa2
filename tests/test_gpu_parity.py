"""GPU parity: the HIP step against the CPU oracle on identical packed worlds.

Integer work (broadphase pair sets, contact records and their canonical order, per-row colors) must match bit
for bit; the kernels keep the oracle's float operation order, so whole trajectories match bit for bit too,
which is stronger than the 1e-5 relative one-step bound (also checked, against the FP64 oracle)."""
import os
import sys

import numpy as np
import pytest

from oracle_bind import OracleWorld, warm_table_map
from steel_webgpu_amd import SceneBuilder, pack, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim():
    from steel_webgpu_amd import sim as s
    s.default_device()
    return s


def driven_quadruped(index=3):
    """Quadruped whose 12 joints hold their initial angles with PD drives; efforts mix unlimited (0), loose and
    tight enough to saturate under the legs' load, and some joints also carry passive springs and damping."""
    scene = workloads.quadruped_world(index, drive=dict(targetKe=60.0, targetKd=2.0, armature=0.01, damping=0.05))
    J = scene.json["joints"]
    n = len(J["ke"])
    J["effort"] = [(0.0, 0.3, 5.0)[k % 3] for k in range(n)]
    J["spring_stiffness"] = [4.0 if k % 4 == 1 else 0.0 for k in range(n)]
    J["spring_ref"] = [0.2 if k % 4 == 1 else 0.0 for k in range(n)]
    return scene


def _with_body_limits(scene):
    """Free-body velocity limits on every other body (1 m/s, 2 rad/s); the rest unbounded (-1)."""
    B = scene.json["bodies"]
    n = len(B["mass"])
    B["max_linear_velocity"] = [1.0 if i % 2 else -1.0 for i in range(n)]
    B["max_angular_velocity"] = [2.0 if i % 2 else -1.0 for i in range(n)]
    return scene


def _with_sleeping(scene, quiet=0.1):
    scene.json["solver"].update(enable_sleeping=True, sleep_quiet_time=quiet)
    return scene


def _with_warm(scene):
    """Contact warm start (the reference's contact_warmstart solver option, scale 1)."""
    scene.json["solver"].update(contact_warmstart=True)
    return scene


def _with_nesterov(scene, warm=False):
    """Nesterov momentum on the contact rows (with contact warm start on top when warm)."""
    scene.json["solver"].update(nesterov_momentum=True, contact_warmstart=warm)
    return scene


def _with_torsion(scene, radius=0.02):
    """Contact torsion radius on every shape (the reference's torsion_radius; articulated contacts use it)."""
    for s in scene.json["shapes"]:
        s["torsion_radius"] = radius
    return scene


def d6_rig():
    """A floating box with three D6-jointed boxes (three rotations; two rotations; two translations and a rotation)
    dropping onto the ground."""
    from steel_webgpu_amd.scene import JT
    X, Y, Z = [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]
    b = SceneBuilder(fps=240)
    b.addGroundPlane()
    base = b.addBody(xform=[0, 0, 0.5, 0, 0, 0, 1])
    b.addShapeBox(base, hx=0.3, hy=0.1, hz=0.05)
    b.addJointFree(base)
    for k, (axes, lin, q) in enumerate([([X, Y, Z], 0, [0.3, 0.2, 0.1]), ([X, Y], 0, [0.4, -0.3]), ([X, Y, Z], 2, [0.05, 0.02, 0.5])]):
        x = -0.25 + 0.25 * k
        c = b.addBody(xform=[x, 0, 0.5, 0, 0, 0, 1])
        b.addShapeBox(c, hx=0.04, hy=0.04, hz=0.1, xform=[0, 0, -0.15, 0, 0, 0, 1])
        b.addJoint(JT["D6"], base, c, parentXform=[x, 0, 0, 0, 0, 0, 1], axes=axes, linear=lin, q=q)
    return b.finalize(solver=dict(substeps=4))


def kinematic_turntable(n=6):
    """A kinematic slab spinning at 1.2 rad/s (prescribed twist, infinite mass, no gravity) carrying boxes."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane()
    t = b.addBody(xform=[0, 0, 0.05, 0, 0, 0, 1], kinematic=True, twist=(0, 0, 0, 0, 0, 1.2))
    b.addShapeBox(t, hx=0.6, hy=0.6, hz=0.05)
    for k in range(n):
        bx = b.addBody(xform=[0.12 * (k - 2.5), 0.03 * k, 0.15 + 0.001 * k, 0, 0, 0, 1])
        b.addShapeBox(bx, hx=0.05, hy=0.05, hz=0.05)
    return b.finalize(solver=dict(substeps=4))


def _with_solver(scene, **opts):
    scene.json["solver"].update(opts)
    return scene


def loop_rig():
    """Two closed chains: a hinged bar plus a bar on a ball joint, each closed to the world at its foot, one by a
    revolute loop joint, one by a ball loop joint; the ball-jointed bars start spinning; a box drops on each."""
    from steel_webgpu_amd.scene import JT
    b = SceneBuilder(fps=240)
    b.addGroundPlane()
    for k, lt in enumerate(("REVOLUTE", "BALL")):
        y = 0.6 * k
        A = b.addBody(xform=[0.25, y, 1, 0, 0, 0, 1])
        b.addShapeBox(A, hx=0.25, hy=0.02, hz=0.02)
        B = b.addBody(xform=[0.5, y, 0.75, 0, 0, 0, 1])
        b.addShapeBox(B, hx=0.02, hy=0.02, hz=0.25)
        b.addJointRevolute(-1, A, axis=[0, 1, 0], parentXform=[0, y, 1, 0, 0, 0, 1], childXform=[-0.25, 0, 0, 0, 0, 0, 1])
        b.addJointBall(A, B, parentXform=[0.25, 0, 0, 0, 0, 0, 1], childXform=[0, 0, 0.25, 0, 0, 0, 1])
        axes = [[0, 1, 0]] if lt == "REVOLUTE" else [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
        b.addJoint(JT[lt], -1, B, parentXform=[0.5, y, 0.5, 0, 0, 0, 1], childXform=[0, 0, -0.25, 0, 0, 0, 1], axes=axes,
                   q=[0.0] if lt == "REVOLUTE" else [0, 0, 0, 1])
        bx = b.addBody(xform=[0.3, y, 1.4, 0, 0, 0, 1])
        b.addShapeBox(bx, hx=0.06, hy=0.06, hz=0.06)
    return b.finalize(solver=dict(substeps=4))


def spin_roll_rig():
    """Balls rolling and spinning with spin and rolling rows (condim 6), and boxes with spin rows (condim 4) dropping
    onto them and the ground."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane()
    for k in range(4):
        s = b.addBody(xform=[0.3 * k, 0, 0.08, 0, 0, 0, 1], twist=(0.2 * k, 0, 0, 0, 2.0 + k, 3.0 - k))
        b.addShapeSphere(s, radius=0.08)
    for k in range(4):
        bx = b.addBody(xform=[0.3 * k + 0.05, 0.02, 0.4 + 0.1 * k, 0, 0, 0, 1], twist=(0, 0, 0, 0, 0, 4.0))
        b.addShapeBox(bx, hx=0.05, hy=0.05, hz=0.05)
    scene = b.finalize(solver=dict(substeps=4))
    for i, sh in enumerate(scene.json["shapes"]):
        sh.update(condim=6 if 1 <= i <= 4 else 4, mu_torsional=0.02, mu_rolling=0.01)
    return scene


def fader_rig():
    """A slider on a tilted rail coupled to a velocity-driven wheel by the fader controller, plus a box on the wheel."""
    b = SceneBuilder(fps=240)
    b.addGroundPlane()
    s = b.addBody(xform=[0, 0, 1, 0, 0, 0, 1])
    b.addShapeBox(s, hx=0.03, hy=0.03, hz=0.03)
    b.addJointPrismatic(-1, s, axis=[1, 0, 0], parentXform=[0, 0, 1, 0, 0.2588190, 0, 0.9659258])
    w = b.addBody(xform=[1, 0, 0.3, 0, 0, 0, 1])
    b.addShapeCylinder(w, radius=0.3, halfHeight=0.02)
    b.addJointRevolute(-1, w, axis=[0, 0, 1], parentXform=[1, 0, 0.3, 0, 0, 0, 1], targetKd=5.0)
    bx = b.addBody(xform=[1.15, 0, 0.4, 0, 0, 0, 1])
    b.addShapeBox(bx, hx=0.04, hy=0.04, hz=0.04)
    scene = b.finalize(solver=dict(substeps=4))
    scene.json["control"] = dict(type="fader", fader_q=0, wheel_v=1, travel=0.24, max_speed=6.283185307179586)
    return scene


def _with_velocity_iterations(scene, n):
    scene.json["solver"].update(velocity_iterations=n)
    return scene


def _with_mimics(scene):
    """Each leg's third joint mimics its second (q = 0.1 - 1.2 q_leader)."""
    J = scene.json["joints"]
    for leg in range(4):
        f, r = 1 + leg * 3 + 2, 1 + leg * 3 + 1
        J["mimic_joint"][f] = r
        J["mimic_coeffs"][f] = [0.1, -1.2]
    return scene


def _with_joint_friction(scene):
    """Load-dependent joint friction (coefficient 0.02) on top of the plain friction bound."""
    J = scene.json["joints"]
    J["friction_load_coefficient"] = [0.02] * len(J["friction"])
    return scene


def big_hull_pile(seed=21, n=27):
    """Hulls of 40 vertices, boxes with proud edge midpoints (20 vertices) and cylinders (48-vertex hulls,
    analytic support), dropped in random orientations onto a static slab and the ground: the MPR path, plane and
    box contacts of cylinders and their rolling."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder(fps=240)
    b.addGroundPlane(mu=0.6)
    b.addShapeBox(-1, hx=0.4, hy=0.4, hz=0.05, xform=[0.3, 0.0, 0.05, 0, 0, 0, 1], mu=0.6)
    for i in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        x, y, z = (i % 3) * 0.3 - 0.2, ((i // 3) % 3) * 0.3 - 0.3, 0.35 + (i // 9) * 0.3
        bd = b.addBody(xform=[x, y, z, *q])
        k = i % 3
        if k == 0:
            v = rng.normal(size=(40, 3))
            v *= rng.uniform(0.05, 0.1) / np.linalg.norm(v, axis=1, keepdims=True)
            b.addShapeConvexHull(bd, v.tolist(), mu=0.6)
        elif k == 1:
            h = rng.uniform(0.04, 0.08)
            pts = [[sx * h, sy * h, sz * h] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
            for ax in range(3):
                o = [j for j in range(3) if j != ax]
                for s1 in (-1, 1):
                    for s2 in (-1, 1):
                        p = [0.0, 0.0, 0.0]
                        p[o[0]], p[o[1]] = s1 * h * 1.01, s2 * h * 1.01
                        pts.append(p)
            b.addShapeConvexHull(bd, pts, mu=0.6)
        else:
            b.addShapeCylinder(bd, radius=rng.uniform(0.04, 0.07), halfHeight=rng.uniform(0.05, 0.12), mu=0.6)
    return b.finalize(solver=dict(substeps=4, iterations=12))


SCENES = {
    "boxes64": lambda: workloads.boxes64(),
    "quad0": lambda: workloads.quadruped_world(0),
    "quad5": lambda: workloads.quadruped_world(5),
    "arm": lambda: workloads.arm_hulls(),
    "quad_driven": driven_quadruped,
    "serve1500": lambda: workloads.serving_request(11, nmin=1500, nmax=1500)[0],
    "bighulls": big_hull_pile,
    "boxes_sleep": lambda: _with_sleeping(workloads.boxes64()),
    "quad_sleep": lambda: _with_sleeping(workloads.quadruped_world(9)),
    "quad_mimic": lambda: _with_mimics(workloads.quadruped_world(10)),
    "boxes_warm": lambda: _with_warm(workloads.boxes64()),
    "quad_warm": lambda: _with_warm(workloads.quadruped_world(12)),
    "boxes_nest": lambda: _with_nesterov(workloads.boxes64()),
    "quad_nest": lambda: _with_nesterov(workloads.quadruped_world(13), warm=True),
    "quad_tors": lambda: _with_torsion(workloads.quadruped_world(14)),
    "d6": d6_rig,
    "boxes_vel": lambda: _with_velocity_iterations(workloads.boxes64(), 4),
    "kinematic": kinematic_turntable,
    "loops": loop_rig,
    "spin_roll": spin_roll_rig,
    "fader": fader_rig,
    "boxes_desaxce": lambda: _with_solver(workloads.boxes64(n=40), friction_mode="bisection_desaxce"),
    "quad_newton": lambda: _with_solver(workloads.quadruped_world(18), friction_mode="coulomb_newton"),
    "quad_bisect": lambda: _with_solver(workloads.quadruped_world(19), friction_mode="bisection"),
    # articulated spin/rolling through the elliptic cone block (anchors off), with torsion-radius rows too
    "quad_cone": lambda: _with_torsion(_with_solver(workloads.quadruped_world(20), enable_torsional_rolling_friction=True)),
    # drive rows with velocity limits (driven dofs fused into the sweep clamp) and effort limits
    "quad_drows": lambda: _with_solver(workloads.quadruped_world(17, drive=dict(targetKe=60.0, targetKd=2.0, velocityLimit=0.8, effortLimit=4.0)),
                                       drive_mode="rows", enable_joint_velocity_limits=True),
    "boxes_compl": lambda: _with_solver(workloads.boxes64(n=40), contact_compliance=True),
    "quad_compl": lambda: _with_solver(workloads.quadruped_world(16), contact_compliance=True),
    "quad_vel": lambda: _with_velocity_iterations(_with_mimics(workloads.quadruped_world(15)), 2),
    # a recorded 60 Hz gait driving the legs (base dofs of the path are 0; their drives are off)
    "quad_path": lambda: workloads.with_target_path(workloads.quadruped_world(8, drive=dict(targetKe=40.0, targetKd=1.0)),
                                                    workloads.gait_path(), integral=8),
    "boxes_limits": lambda: _with_body_limits(workloads.boxes64(n=40)),
    "quad_friction": lambda: _with_joint_friction(workloads.quadruped_world(6, drive=dict(targetKe=30.0, targetKd=1.0, friction=0.05),
                                                                          solver=dict(enable_joint_friction=True))),
    # joint velocity limits of 0.5 rad/s on driven legs (rows active at any speed: activation fraction 0)
    "quad_vlim": lambda: workloads.quadruped_world(4, drive=dict(targetKe=60.0, targetKd=2.0, velocityLimit=0.5),
                                                   solver=dict(enable_joint_velocity_limits=True)),
}


sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import scene_worlds  # noqa: E402


@pytest.mark.parametrize("name", scene_worlds.names())
def test_reference_scene_folders_bitexact(sim, name):
    """Each of the reference's 41 shipped scene folders (packed fixtures): 150 frames from their initial state,
    GPU equal to the FP32 oracle bit for bit (poses, velocities, articulation state, stats words)."""
    pw = scene_worlds.load(name)
    b = sim.Batch([pw])
    o = OracleWorld(pw)
    b.step(150)
    o.frames(150)
    assert np.array_equal(_bits(b.read(0, "poses")), _bits(o.poses))
    assert np.array_equal(_bits(b.read(0, "vel")), _bits(o.vel[:2 * max(pw.NB, 1) * 4]))
    assert np.array_equal(_bits(b.read(0, "art")), _bits(o.art))
    assert np.array_equal(b.stats_raw(0)[0][:9], o.stats[:9])
    b.destroy()


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", ["boxes64", "quad5", "arm", "serve1500", "bighulls"])
def test_broadphase_pair_set_and_contacts_bitexact(sim, name):
    pw = pack.pack_scene(SCENES[name]())
    b = sim.Batch([pw], graph=False)
    o = OracleWorld(pw)
    # advance both a few frames so the pile is in contact, then compare one isolated collide stage
    b.step(30)
    o.frames(30)
    b.substep_stages(sim.STAGE_COLLIDE | sim.STAGE_NARROW)
    o.substeps(1, debug=True)
    ctr = b.read(0, "counters", dtype=np.uint32)
    gp = b.read(0, "pairs", dtype=np.uint32)[:2 * min(ctr[0], pw.h("W_MAXP"))].reshape(-1, 2)
    op = o.dbg_bufs["pairs"][:2 * int(o.dbg_bufs["npairs"][0])].reshape(-1, 2)
    assert ctr[0] == o.dbg_bufs["npairs"][0]
    assert set(map(tuple, gp.tolist())) == set(map(tuple, op.tolist()))
    nc = int(o.dbg_bufs["ncontacts"][0])
    assert ctr[1] == nc and nc > 0
    gc = b.read(0, "contacts", nbytes=nc * 96, dtype=np.float32)
    assert np.array_equal(_bits(gc), _bits(o.dbg_bufs["contacts"][:nc * 24]))
    b.destroy()


@pytest.mark.parametrize("name,frames", [("boxes64", 120), ("quad0", 200), ("quad5", 80), ("arm", 12), ("quad_driven", 200), ("serve1500", 40), ("quad_vlim", 200),
                                         ("boxes_limits", 150), ("quad_friction", 200), ("bighulls", 120),
                                         ("quad_path", 200), ("boxes_sleep", 400), ("quad_sleep", 300),
                                         ("quad_mimic", 200), ("boxes_warm", 150), ("quad_warm", 200),
                                         ("boxes_nest", 150), ("quad_nest", 200), ("quad_tors", 200), ("d6", 200),
                                         ("boxes_vel", 150), ("quad_vel", 200), ("kinematic", 200),
                                         ("boxes_compl", 150), ("quad_compl", 200), ("quad_drows", 200),
                                         ("loops", 200), ("spin_roll", 200), ("fader", 200),
                                         ("boxes_desaxce", 150), ("quad_newton", 200), ("quad_bisect", 200),
                                         ("quad_cone", 200)])
def test_trajectory_bitexact(sim, name, frames):
    pw = pack.pack_scene(SCENES[name]())
    b = sim.Batch([pw])
    o = OracleWorld(pw)
    b.step(frames)
    o.frames(frames)
    assert np.array_equal(_bits(b.read(0, "poses")), _bits(o.poses))
    assert np.array_equal(_bits(b.read(0, "vel")), _bits(o.vel[:2 * max(pw.NB, 1) * 4]))
    assert np.array_equal(_bits(b.read(0, "art")), _bits(o.art))
    gs = b.stats_raw(0)[0]
    assert np.array_equal(gs[:9], o.stats[:9])
    if name.endswith("sleep"):
        assert o.stats[8] > 0  # some islands are asleep
    if name.endswith("warm"):
        # both parities of the warm-start table hold the same key -> impulse entries (slots may differ)
        WH = pw.h("W_WH")
        gw, ow = b.read(0, "warm", dtype=np.uint32), o.warm_u32()
        for par in (0, 1):
            gm, om = warm_table_map(gw, WH, par), warm_table_map(ow, WH, par)
            assert gm.keys() == om.keys() and len(om) > 0
            assert all(np.array_equal(_bits(gm[k]), _bits(om[k])) for k in om)
        assert o.stats[7] > 0 or name.startswith("boxes")  # dense rows stored and read back
    b.destroy()


@pytest.mark.parametrize("name", ["boxes64", "quad5"])
def test_coloring_bitexact(sim, name):
    pw = pack.pack_scene(SCENES[name]())
    b = sim.Batch([pw])
    o = OracleWorld(pw)
    b.step(120)
    o.frames(119)
    o.substeps(pw.h("W_SUBSTEPS") - 1)
    o.substeps(1, debug=True)
    L = pack.world_sched_layout(pw)
    sched = b.read(0, "sched", dtype=np.uint32)
    nr = int(o.dbg_bufs["nrows"][0])
    info = o.dbg_bufs["rowinfo"][:nr * 4].reshape(nr, 4)
    assert nr > 20
    assert np.array_equal(sched[L["PCOL"]:L["PCOL"] + nr], info[:, 1])
    assert np.array_equal(sched[L["RUNIT"]:L["RUNIT"] + nr], info[:, 2])
    assert np.array_equal(sched[L["META"]:L["META"] + 4], o.dbg_bufs["meta"][:4])
    b.destroy()


def test_worlds_beyond_the_lds_capacities_take_the_global_paths_bitexact(sim):
    """5640 shapes exceed k_broad's LDS sweep (4096 shapes) and 5639 bodies k_prepare's LDS coloring masks
    (16384 words at 3 per body): both run from global memory, with the same results, in a batch beside a small
    world that keeps the LDS paths."""
    big = pack.pack_scene(workloads.serving_request(11, nmin=5600, nmax=5600)[0], capacity_policy="compact")
    small = pack.pack_scene(workloads.quadruped_world(3))
    assert big.NS > 4096 and big.NB * 3 > 16384
    b = sim.Batch([small, big])
    ob, osm = OracleWorld(big), OracleWorld(small)
    b.step(16)
    ob.frames(16)
    osm.frames(16)
    for i, o in ((0, osm), (1, ob)):
        assert np.array_equal(_bits(b.read(i, "poses")), _bits(o.poses)), i
        assert np.array_equal(_bits(b.read(i, "art")), _bits(o.art)), i
    assert b.stats_raw(1)[0][5] > 0
    b.destroy()


def test_batched_equals_alone_and_mixed_sizes(sim):
    """A world's result does not depend on the other worlds of its batch (alone vs mixed, bitwise)."""
    worlds = [pack.pack_scene(workloads.quadruped_world(i), capacity_policy="compact") for i in (0, 3, 5, 11, 17)]
    worlds.insert(2, pack.pack_scene(workloads.boxes64(n=24)))
    mixed = sim.Batch(worlds)
    mixed.step(60)
    for i, pw in enumerate(worlds):
        alone = sim.Batch([pw])
        alone.step(60)
        for name in ("poses", "vel", "art"):
            assert np.array_equal(_bits(mixed.read(i, name)), _bits(alone.read(0, name))), (i, name)
        alone.destroy()
    mixed.destroy()


@pytest.mark.parametrize("name", ["quad5", "boxes_nest"])
def test_deterministic_across_runs_block_sizes_and_graphs(sim, name):
    pw = pack.pack_scene(SCENES[name]())
    ref = None
    for block, graph in ((256, True), (256, True), (64, True), (128, False), (64, False)):
        b = sim.Batch([pw], block=block, graph=graph)
        b.step(50)
        got = [_bits(b.read(0, n)) for n in ("poses", "vel", "art")]
        if ref is None:
            ref = got
        for g, r in zip(got, ref):
            assert np.array_equal(g, r), (block, graph)
        b.destroy()


def test_reset_restores_initial_state(sim):
    pw = pack.pack_scene(workloads.boxes64(n=20))
    b = sim.Batch([pw])
    p0 = b.read(0, "poses").copy()
    b.step(30)
    p1 = b.read(0, "poses").copy()
    b.reset(0)
    assert np.array_equal(_bits(b.read(0, "poses")), _bits(p0))
    b.step(30)
    assert np.array_equal(_bits(b.read(0, "poses")), _bits(p1))
    b.destroy()


def test_joint_targets_set_mid_run_bitexact(sim):
    """Sim.setJointTargets (a policy's action vector, sim.js:716-728) between frames: the GPU world and the oracle,
    given the same targets, stay bitwise equal; the drives then pull the joints toward the new targets."""
    from steel_webgpu_amd.sim import Sim
    scene = driven_quadruped(7)
    s = Sim(None, scene)
    pw = s.packed
    o = OracleWorld(pw)
    A, ND = pw.A, pw.ND
    rng = np.random.default_rng(3)
    J = scene.json["joints"]
    first = J["qd_start"][1]        # Newton dof of the first leg joint (joint 0 is the floating base)
    for chunk in range(3):
        tgt = rng.uniform(-0.4, 0.4, 12).astype(np.float32)
        tv = rng.uniform(-0.5, 0.5, 12).astype(np.float32)
        s.setJointTargets(tgt, tv, first)
        for k in range(12):
            d = pw.dofMap[first + k]
            o.art[A["A_TGT"] + d] = tgt[k]
            o.art[A["A_TGTV"] + d] = tv[k]
        s.step(40)
        o.frames(40)
        assert np.array_equal(_bits(s.batch.read(0, "art")), _bits(o.art)), chunk
        assert np.array_equal(_bits(s.batch.read(0, "poses")), _bits(o.poses)), chunk
    assert s.userTargets


def test_one_step_error_against_fp64_within_1e5(sim):
    """One GPU frame from the initial state against the FP64 oracle's frame, per quantity (largest difference over
    the largest FP64 magnitude): body positions and orientations and joint coordinates within 1e-5; body twists
    and joint velocities within 1e-3 (contact rows that start or stop between the float and double solves move
    velocities by more than rounding; the GPU equals the FP32 oracle bit for bit, so this is FP32's own error)."""
    from oracle_bind import state_quantities
    for name in ("boxes64", "quad0", "quad5", "quad_driven", "quad_vlim", "boxes_limits", "quad_friction", "bighulls", "quad_path",
                 "quad_mimic"):
        pw = pack.pack_scene(SCENES[name]())
        b = sim.Batch([pw])
        o64 = OracleWorld(pw, f64=True)
        b.step(1)
        o64.frames(1)
        g = state_quantities(pw, b.read(0, "poses"), b.read(0, "vel"), b.read(0, "art"))
        r = state_quantities(pw, o64.poses, o64.vel, o64.art)
        for q in g:
            if r[q].size == 0 or np.abs(r[q]).max() == 0:
                continue
            rel = np.abs(g[q] - r[q]).max() / np.abs(r[q]).max()
            bound = 1e-5 if q in ("position", "orientation", "q") else 1e-3
            assert rel <= bound, (name, q, rel)
        b.destroy()


def test_arm_among_hulls_from_a_settled_checkpoint_bitexact_with_dense_rows(sim):
    """The arm among 256 hulls from the FP32 oracle's frame-400 state (committed checkpoint: hulls settled on the
    floor and on the arm): 20 more frames on the GPU and on the oracle match bit for bit, state, stats and anchors,
    and the arm's self-contacts run as dense rows in that window."""
    import make_checkpoints
    pw = make_checkpoints.world("arm_hulls")
    o = OracleWorld(pw)
    frame = make_checkpoints.load_into(o, "arm_hulls")
    assert frame == 400
    o.stats[7] = 0   # the window's largest dense-row count, on both sides
    b = sim.Batch([pw], block=64)
    b.write(0, "poses", o.poses)
    b.write(0, "vel", o.vel[:2 * max(pw.NB, 1) * 4])
    b.write(0, "art", o.art)
    b.write(0, "anchor_keys", o.akey)
    b.write(0, "anchor_vals", o.aval)  # the oracle's float anchor record has the kernels' 14-word layout
    ctr = np.zeros(8, np.uint32)
    ctr[2], ctr[3], ctr[4], ctr[5] = o.count[0], o.count[1], o.count[2], o.count[3]
    b.write(0, "counters", ctr)
    b.write(0, "stats", o.stats)
    for _ in range(4):
        b.step(5)
        o.frames(5)
        assert np.array_equal(_bits(b.read(0, "poses")), _bits(o.poses))
        assert np.array_equal(_bits(b.read(0, "vel")), _bits(o.vel[:2 * max(pw.NB, 1) * 4]))
        assert np.array_equal(_bits(b.read(0, "art")), _bits(o.art))
    gs = b.stats_raw(0)[0]
    assert np.array_equal(gs[:9], o.stats[:9])
    assert o.stats[7] > 0, "no dense rows in the window"
    # anchor tables as key -> record maps (slots are taken by atomic insertion on the GPU, serially on the CPU)
    AH = pw.h("W_AH")
    gk, gv = b.read(0, "anchor_keys", dtype=np.uint32), b.read(0, "anchor_vals", dtype=np.uint32).reshape(-1, 14)
    ot = o.anchors_u32()
    ok, ov = ot[:2 * AH], ot[2 * AH:].reshape(-1, 14)
    gmap = {(i // AH, int(k)): gv[i].tobytes() for i, k in enumerate(gk) if k}
    omap = {(i // AH, int(k)): ov[i].tobytes() for i, k in enumerate(ok) if k}
    assert gmap == omap and len(omap) > 0
    b.destroy()


def test_contact_overflow_reported_and_bitexact_until_it_happens(sim):
    """Contacts beyond a world's capacity are dropped and reported, as in the reference, whose narrowphase takes
    slots with atomicAdd and drops those at or beyond MAXC (src/wgsl.js:293-296 emit; :3880 stats word 1 = max requested count).
    Which contacts survive depends on emission order (atomic order there and here), so bitwise parity with the
    oracle holds up to the first overflowing substep; there the kept records are exactly MAXC records of the
    full-capacity contact set, and the state stays finite afterwards."""
    CAP = 80
    small = pack.pack_scene(workloads.boxes64(), capacity_override={"MAXC": CAP})
    full = pack.pack_scene(workloads.boxes64())
    g = sim.Batch([small], graph=False)
    o = OracleWorld(small)
    S = small.h("W_SUBSTEPS")
    hit = False
    for sub in range(300 * S):
        pre = g.read(0, "poses").copy()
        g.substep_stages(sim.STAGE_COLLIDE | sim.STAGE_NARROW)
        requested = int(g.read(0, "counters", dtype=np.uint32)[1])
        if requested > CAP:
            kept = g.read(0, "contacts", nbytes=CAP * 96, dtype=np.uint32).reshape(CAP, 24)
            fb = sim.Batch([full], graph=False)
            fb.write(0, "poses", pre)
            fb.substep_stages(sim.STAGE_COLLIDE | sim.STAGE_NARROW)
            nfull = int(fb.read(0, "counters", dtype=np.uint32)[1])
            assert nfull == requested
            allrec = fb.read(0, "contacts", nbytes=nfull * 96, dtype=np.uint32).reshape(nfull, 24)
            fb.destroy()
            pool = set(map(bytes, allrec))
            assert all(bytes(r) in pool for r in kept)
            assert len(set(map(bytes, kept))) == CAP
            hit = True
            break
        g.substep_stages(sim.STAGE_PREPARE | sim.STAGE_SOLVE)
        o.substeps(1)
        assert np.array_equal(_bits(g.read(0, "poses")), _bits(o.poses)), sub
    assert hit, "the pile never exceeded the contact capacity"
    g.substep_stages(sim.STAGE_PREPARE | sim.STAGE_SOLVE)
    g.step(60)
    assert np.all(np.isfinite(g.read(0, "poses")))
    assert g.stats_raw(0)[0][1] >= CAP
    g.destroy()


def test_sim_api_state_layout(sim):
    from steel_webgpu_amd.sim import Sim
    scene = workloads.quadruped_world(2)
    s = Sim(None, scene)
    s.step(5)
    st = s.state()
    J = scene.json["joints"]
    assert st["bodyQ"].shape == (s.NB * 7,) and st["bodyQd"].shape == (s.NB * 6,)
    assert st["jointQ"].shape == (len(J["q"]),) and st["jointQd"].shape == (J["qd_start"][-1],)
    q = st["bodyQ"].reshape(-1, 7)[:, 3:]
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-5)
    stats = s.stats()
    assert stats["overflow"] == [] and stats["contacts"] >= 0
    s.destroy()
