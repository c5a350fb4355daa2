/* Packed world format: the data contract between the host packer and the step kernels.
 *
 * One world is a set of flat arrays with the record layouts of the reference's buffers
 * (packages/steel-webgpu/src/sim.js:212-237 statF/statU sections, :438 poses, :452-455 art, :5-22 names):
 *   statF  float4[W_NSTATF]   body / shape / hull / link / dof records (read only)
 *   statU  uint32[W_NSTATU]   hull topology, articulation table, collision-filter bitset (read only)
 *   poses  float4[W_NPOSES]   body COM pose, shape poses, shape AABBs, previous body poses
 *   vel    float4[2*NB]       free-body COM twists (v, 0), (w, 0)
 *   art    float [W_NART]     articulation state and per-step factor blocks
 *   warm   uint32[W_NWARM]    friction-anchor tables (two parities)
 * plus a header of uint32 words, indexed by the enum below (float fields hold float bits).
 * Capacities (MAXP, MAXC, ...) are per world; a world that overflows one drops work and reports it in
 * its stats, as the reference does (sim.js:826-828).
 */
#ifndef STEEL_WORLD_H
#define STEEL_WORLD_H

#ifdef __cplusplus
extern "C" {
#endif

enum steel_world_field {
  /* counts */
  W_NB = 0, W_NS, W_NH, W_NA, W_NL, W_ND, W_NQ, W_NDA, W_JM, W_MAXPATH, W_MAXLP,
  /* capacities */
  W_MAXP, W_MAXC, W_MAXU, W_UH, W_AH, W_MAXDENSE, W_NCOLW,
  /* row record shape and switches */
  W_ROWS, W_PO, W_PATCH, W_EXTRA, W_FRICTION_COMBINE, W_RESTITUTION_COMBINE, W_REST,
  /* statF sections (float4 index) */
  W_SF_BODY, W_SF_SHAPE, W_SF_HAABB, W_SF_VERTS, W_SF_FPLANE, W_SF_LINK, W_SF_DOF, W_NSTATF,
  /* statU sections (uint32 index) */
  W_SU_HULL, W_SU_FACE, W_SU_FV, W_SU_HEDGE, W_SU_EDGE, W_SU_ART, W_SU_FILTER, W_NSTATU,
  /* poses sections (float4 index) */
  W_P_BODY, W_P_SHAPE, W_P_AABB, W_P_PREV, W_NPOSES,
  /* art sections (float index) */
  W_A_Q, W_A_QD, W_A_QDH, W_A_TGT, W_A_TGTV, W_A_LIM, W_A_DRV, W_A_JF, W_A_BF, W_A_FRIC, W_A_HINV,
  W_A_LINK, W_A_LG, W_A_RL, W_A_RLN, W_A_ULOAD, W_A_DENSE, W_NART,
  W_DHDR, W_DPS, W_NWARM,
  /* solver constants (float bits) */
  W_F_CFM, W_F_OMEGA, W_F_SPEC, W_F_GAP_GATE, W_F_FGAP, W_F_FSCALE, W_F_ANCHOR_BETA, W_F_BP_CELL,
  W_F_REST_SPEED, W_F_JL_GAP, W_F_JL_SPEED,
  /* per-substep parameters (float bits unless noted) */
  W_DT, W_BETA, W_MAX_DEPEN, W_GX, W_GY, W_GZ, W_WARM_SCALE, W_REG, W_ITERS /* u32 */, W_SUBSTEPS /* u32 */,
  W_FRAME_DT,
  /* joint velocity limits: switch (u32), activation fraction (float bits), art offsets of the per-dof and
     per-articulation activity flags */
  W_VLIM_ON, W_F_VLIM_FRAC, W_A_VLG, W_A_AVL,
  /* free-body velocity limits (u32 switch); load-dependent joint friction (u32 switch) and its art offset */
  W_VLIM_FREE, W_JFL, W_A_JFL,
  /* recorded target paths: statF section (float4 index), driven dofs, path frames (60 Hz), dofs with the
     integral term, hold-per-frame switch (u32 each), frames per second (float bits), art offset of the
     integral terms */
  W_SF_PATH, W_DRIVEN, W_PATH_FRAMES, W_INTEGRAL_COUNT, W_TGT_HELD, W_FPS, W_A_INT,
  /* sleeping: switch (u32), linear and angular speed thresholds, quiet time (float bits), art offset of the
     per-node sleep records (age, asleep, -, -; free body b is node b, articulation a node NB + a) */
  W_SLEEP, W_F_SLEEP_LIN, W_F_SLEEP_ANG, W_F_SLEEP_TIME, W_A_SLP,
  /* mimic joints: count (u32), statF section of their records (2 float4 each: articulation, follower dof,
     leader dof, c1; follower q, leader q, c0, -), statU section of the per-articulation "owns a constraint" flags */
  W_NMIMIC, W_SF_MIMIC, W_SU_ACON,
  /* contact warm start: slots of the warm table per parity (u32; 0 when warm start is off) */
  W_WH,
  /* Nesterov momentum on the contact rows (u32 flag) and its row-record offset (u32) */
  W_NEST, W_MO,
  /* contact torsion of articulated units (u32 flag) and its row-record offset (u32) */
  W_TORS, W_TO,
  /* velocity iterations after the position sweeps (u32) and the art offset of their velocity snapshot (u32) */
  W_VELITERS, W_A_SNAP,
  /* contact compliance (u32 flag: spring-damper normals, contacts kept whole) and its row-record offset (u32) */
  W_COMPL, W_CO,
  /* drive_mode rows: drives as velocity rows in the sweeps instead of the implicit factor term (u32), velocity
     limits of driven dofs left to the sweep clamp (u32), drive rows active in the velocity pass (u32) */
  W_DRIVE_ROWS, W_FUSE_VLIM, W_VEL_DRIVE_ACTIVE,
  /* loop closures (ball or revolute joints closing inside one articulation or to the world): count (u32) and
     statF section (u32, 5 vec4 each: parent link, child link, articulation, revolute flag; parent and child
     anchors; parent and child hinge axes) */
  W_NLOOP, W_SF_LOOP,
  /* spin and rolling rows of free-body contacts (shapes with condim >= 4; u32 flag; rows 19-33, anchors at 34) */
  W_ANG,
  /* demo controller (the scene's "control"): kind (u32; 1 fader: the wheel dof's target velocity follows the
     fader coordinate, max_speed * clamp(q / travel, 0, 1)), fader coordinate (u32), wheel dof (u32), statF record
     (u32: travel, max_speed) */
  W_CTRL_KIND, W_CTRL_A, W_CTRL_B, W_SF_CTRL,
  /* friction_mode (u32: 0 fixed_normal, 1 bisection, 2 bisection_desaxce, 3 coulomb_newton) and the row-record
     offset of the coupled block's Delassus entries (u32) */
  W_FMODE, W_FMO,
  /* enable_torsional_rolling_friction: articulated contacts solve (t1, t2, spin, roll1, roll2) in Steel's elliptic
     cone block (u32 flag), the row offset of its 5x5 response (u32), the creep speed (f32) */
  W_ANG_ART, W_EO, W_F_CREEP,
  /* collision-filter format (u32): 0 the reference's NS x NS bitset at SU_FILTER (bit i * NS + j); 1 a sorted list per
     shape, for worlds whose bitset would be large: offsets[NS + 1], then for each shape i the shapes j > i it must not
     pair with, ascending */
  W_FILT_LIST,
  /* cell broadphase (f32; 0: the min-x sweep): shapes whose AABB extent is at most this on every axis are binned
     in a uniform grid of cells sorted by Morton code, the rest are tested against every shape; the pair set is the
     reference's either way */
  W_CELL,
  W_FIELD_COUNT
};

#define STEEL_WORLD_HEADER_WORDS 160
#define STEEL_LB 268      /* floats per link block in art (artwgsl.js:8) */
#define STEEL_AVN 14      /* u32 per friction-anchor value record (wgsl.js:96) */

#ifdef __cplusplus
}
#endif
#endif
