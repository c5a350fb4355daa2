'use strict';
// ES2019 wrapper of the Node-API addon (steel_napi.node) for the reference's JS host: the calls sim.js makes on
// WebGPU (packages/steel-webgpu/src/sim.js:63-835, device.js:1-27) mapped onto libsteel_amd.so, plus
// worldHeader(), which writes the 160-word world header from the layout and constants the Sim constructor already
// computes (sim.js:212-488: SF, SU, P, A sections and the C constants).
const fs = require('fs');
const path = require('path');

const addon = require(process.env.STEEL_NAPI || path.join(__dirname, '..', 'steel_napi.node'));

// The reference's error for a missing GPU capability (src/device.js:1-4).
class UnsupportedError extends Error {
  constructor(message) {
    super(message);
    this.name = 'UnsupportedError';
  }
}

function call(fn, ...args) {
  try {
    return fn(...args);
  } catch (e) {
    if (e && e.code === 'STEEL_UNSUPPORTED') throw new UnsupportedError(e.message);
    throw e;
  }
}

// Buffer ids of steel_batch_read / steel_batch_write (include/steel_amd.h).
const BUF = { poses: 0, vel: 1, art: 2, anchorKeys: 3, anchorVals: 4, pairs: 5, contacts: 6, rows: 7, sched: 8, counters: 9,
  stats: 10, warm: 11 };

// initDevice() (src/device.js:6-27): throws UnsupportedError without a usable MI355X.
function initDevice(ordinal = 0) {
  const handle = call(addon.deviceOpen, ordinal);
  return { handle, info: addon.deviceInfo(handle), close: () => addon.deviceClose(handle) };
}

// N packed worlds stepped together (one world: the plain Sim).  A world is
// { header: Uint32Array(160), statF: Float32Array, statU: Uint32Array, poses0, vel0, art0: Float32Array }.
class Batch {
  constructor(device, worlds) {
    this.handle = call(addon.batchCreate, device.handle, worlds);
    this.n = worlds.length;
  }
  step(frames = 1) { call(addon.batchStep, this.handle, frames); }
  reset(world = -1) { call(addon.batchReset, this.handle, world); }
  sync() { call(addon.batchSync, this.handle); }
  bufferSize(world, name) { return call(addon.batchBufferSize, this.handle, world, BUF[name]); }
  read(world, name, dst, byteOffset = 0) { call(addon.batchRead, this.handle, world, BUF[name], byteOffset, dst); return dst; }
  write(world, name, src, byteOffset = 0) { call(addon.batchWrite, this.handle, world, BUF[name], byteOffset, src); }
  stats(world = 0) {
    const words = new Uint32Array(16);
    const gpuMs = call(addon.batchStats, this.handle, world, words);
    return { words, gpuMs };
  }
  destroy() { if (this.handle) { addon.batchDestroy(this.handle); this.handle = null; } }
}

// Field order of the header (enum steel_world_field in include/steel_world.h).
function headerFields(headerPath = path.join(__dirname, '..', '..', 'include', 'steel_world.h')) {
  const txt = fs.readFileSync(headerPath, 'utf8');
  const body = txt.slice(txt.indexOf('enum steel_world_field'), txt.indexOf('W_FIELD_COUNT')).replace(/\/\*[\s\S]*?\*\//g, '');
  const names = body.match(/\bW_[A-Z0-9_]+\b/g);
  const out = {};
  names.forEach((n, i) => { out[n] = i; });
  return out;
}

const f32 = new Float32Array(1);
const u32 = new Uint32Array(f32.buffer);
function fbits(x) { f32[0] = x; return u32[0]; }

// The world header from what the Sim constructor computes (sim.js:212-488): C (this.C, which also spreads the SF,
// SU, P, Sd and A offsets), the section sizes it allocates, and the scene-level values C does not carry.
//   extra: { NH, NQ, nstatf, nstatu, nart, fps, substeps, iterations, gravity: [x, y, z], regularization,
//            warmScale, maxDepen, baumgarte, warmSlots (WH when warm start is on, else 0), capacities
//            { MAXP, MAXC, MAXU, UH, AH, MAXDENSE } (the reference's: 16384, 8192, 2048, 4096, 4096, 512), path:
//            { driven, frames, integral } }
function worldHeader(C, extra, fields = headerFields()) {
  const h = new Uint32Array(160);
  const set = (k, v) => { if (!(k in fields)) throw new Error(`no header field ${k}`); h[fields[k]] = v >>> 0; };
  const setf = (k, v) => { if (!(k in fields)) throw new Error(`no header field ${k}`); h[fields[k]] = fbits(v); };
  const cap = extra.capacities;
  for (const k of ['NB', 'NS', 'NA', 'NL', 'ND', 'NDA', 'JM', 'MAXPATH', 'MAXLP', 'NCOLW', 'ROWS', 'PO', 'PATCH', 'EXTRA', 'REST',
    'DHDR', 'DPS', 'NMIMIC', 'NEST', 'MO', 'TORS', 'TO', 'VELITERS', 'COMPL', 'CO', 'DRIVE_ROWS', 'FUSE_VLIM', 'VEL_DRIVE_ACTIVE',
    'NLOOP', 'ANG', 'CTRL_KIND', 'CTRL_A', 'CTRL_B', 'FMODE', 'FMO', 'ANG_ART', 'EO', 'VLIM_ON', 'VLIM_FREE', 'JFL', 'TGT_HELD',
    'SLEEP']) set('W_' + k, C[k]);
  set('W_FRICTION_COMBINE', C.FRICTION_COMBINE);
  set('W_RESTITUTION_COMBINE', C.RESTITUTION_COMBINE);
  set('W_NH', extra.NH);
  set('W_NQ', extra.NQ);
  for (const k of ['MAXP', 'MAXC', 'MAXU', 'UH', 'AH', 'MAXDENSE']) set('W_' + k, cap[k]);
  for (const k of ['SF_BODY', 'SF_SHAPE', 'SF_HAABB', 'SF_VERTS', 'SF_FPLANE', 'SF_LINK', 'SF_DOF', 'SF_PATH', 'SF_LOOP', 'SF_MIMIC',
    'SF_CTRL', 'SU_HULL', 'SU_FACE', 'SU_FV', 'SU_HEDGE', 'SU_EDGE', 'SU_ART', 'SU_ACON', 'SU_FILTER', 'P_BODY', 'P_SHAPE', 'P_AABB',
    'P_PREV', 'A_Q', 'A_QD', 'A_QDH', 'A_TGT', 'A_TGTV', 'A_LIM', 'A_DRV', 'A_JF', 'A_BF', 'A_FRIC', 'A_HINV', 'A_LINK', 'A_LG',
    'A_RL', 'A_RLN', 'A_ULOAD', 'A_DENSE', 'A_VLG', 'A_AVL', 'A_JFL', 'A_INT', 'A_SLP', 'A_SNAP']) set('W_' + k, C[k]);
  set('W_NSTATF', extra.nstatf);
  set('W_NSTATU', extra.nstatu);
  set('W_NPOSES', C.NB * 4 + C.NS * 4);
  set('W_NART', extra.nart);
  set('W_NWARM', 2 * cap.AH + 2 * cap.AH * 14);
  set('W_ITERS', extra.iterations);
  set('W_SUBSTEPS', extra.substeps);
  set('W_WH', extra.warmSlots);
  set('W_DRIVEN', extra.path.driven);
  set('W_PATH_FRAMES', extra.path.frames);
  set('W_INTEGRAL_COUNT', extra.path.integral);
  for (const k of ['F_CFM', 'F_OMEGA', 'F_SPEC', 'F_GAP_GATE', 'F_FGAP', 'F_FSCALE', 'F_ANCHOR_BETA', 'F_BP_CELL', 'F_REST_SPEED',
    'F_JL_GAP', 'F_JL_SPEED', 'F_VLIM_FRAC', 'F_SLEEP_LIN', 'F_SLEEP_ANG', 'F_SLEEP_TIME', 'F_CREEP']) setf('W_' + k, C[k]);
  const dt = 1 / extra.fps / extra.substeps;
  setf('W_DT', dt);
  setf('W_BETA', extra.baumgarte);
  setf('W_MAX_DEPEN', extra.maxDepen);
  setf('W_GX', extra.gravity[0]);
  setf('W_GY', extra.gravity[1]);
  setf('W_GZ', extra.gravity[2]);
  setf('W_WARM_SCALE', extra.warmScale);
  setf('W_REG', extra.regularization);
  setf('W_FRAME_DT', 1 / extra.fps);
  setf('W_FPS', extra.fps);
  return h;
}

module.exports = { UnsupportedError, BUF, initDevice, Batch, worldHeader, headerFields, lastError: addon.lastError };
