"""ctypes binding of the CPU oracle (oracle/liboracle.so); test infrastructure."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")


class Dbg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("pairs", "npairs", "contacts", "ncontacts", "rowinfo", "nrows", "meta", "rows", "red", "redf")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.dirname(ORACLE)])
        L = ctypes.CDLL(ORACLE)
        L.oracle_anchor_val_size_f32.restype = ctypes.c_size_t
        L.oracle_anchor_val_size_f64.restype = ctypes.c_size_t
        L.oracle_warm_slot_size_f32.restype = ctypes.c_size_t
        L.oracle_warm_slot_size_f64.restype = ctypes.c_size_t
        L.oracle_hash_u32.restype = ctypes.c_uint32
        L.oracle_hash_u32.argtypes = [ctypes.c_uint32]
        for f in ("oracle_sinf", "oracle_cosf"):
            getattr(L, f).restype = ctypes.c_float
            getattr(L, f).argtypes = [ctypes.c_float]
        L.oracle_atan2f.restype = ctypes.c_float
        L.oracle_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_friction_pair_solve.argtypes = [ctypes.c_float] * 9 + [ctypes.c_void_p]
        L.oracle_mpr_f64.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] * 2 + [ctypes.c_double, ctypes.c_void_p]
        L.oracle_gsup_f64.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint32] + [ctypes.c_void_p] * 2
        _lib = L
    return _lib


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class WorldPtrs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("header", "statF", "statU", "poses", "vel", "art", "akey", "aval", "warm", "clock", "stats")]


class OracleWorld:
    """One packed world stepped by the CPU oracle (float = parity oracle, double = error reference)."""

    def __init__(self, pw, f64=False):
        L = lib()
        self.pw = pw
        self.f64 = f64
        dt = np.float64 if f64 else np.float32
        self.poses = pw.poses0.astype(dt)
        self.vel = (pw.vel0 if pw.vel0.size else np.zeros(8, np.float32)).astype(dt)
        self.art = pw.art0.astype(dt)
        AH = pw.h("W_AH")
        self.akey = np.zeros(2 * AH, np.uint32)
        vs = L.oracle_anchor_val_size_f64() if f64 else L.oracle_anchor_val_size_f32()
        self.aval = np.zeros(2 * AH * vs, np.uint8)
        ws = L.oracle_warm_slot_size_f64() if f64 else L.oracle_warm_slot_size_f32()
        self.warm = np.zeros(max(2 * pw.h("W_WH"), 1) * ws, np.uint8)
        self.count = np.zeros(4, np.uint32)  # substep count, control flags, frame clock (double)
        self.stats = np.zeros(16, np.uint32)
        self.statU = pw.statU if pw.statU.size else np.zeros(1, np.uint32)
        MAXC, MAXP = pw.h("W_MAXC"), pw.h("W_MAXP")
        self.dbg_bufs = dict(pairs=np.zeros(2 * MAXP, np.uint32), npairs=np.zeros(1, np.uint32), contacts=np.zeros(MAXC * 24, np.float32),
                             ncontacts=np.zeros(1, np.uint32), rowinfo=np.zeros(MAXC * 4, np.uint32), nrows=np.zeros(1, np.uint32),
                             meta=np.zeros(8, np.uint32), rows=np.zeros(MAXC * pw.h("W_ROWS") * 4, np.float32),
                             red=np.zeros(MAXC * 8, np.uint32), redf=np.zeros(MAXC * 4, np.float32))
        self.dbg = Dbg(**{k: P(v).value for k, v in self.dbg_bufs.items()})

    def substeps(self, n, debug=False):
        L = lib()
        fn = L.oracle_substeps_f64 if self.f64 else L.oracle_substeps_f32
        rc = fn(P(self.pw.header), P(self.pw.statF), P(self.statU), P(self.poses), P(self.vel), P(self.art), P(self.akey), P(self.aval),
                P(self.warm), P(self.count), ctypes.c_uint32(n), P(self.stats), ctypes.byref(self.dbg) if debug else None)
        assert rc == 0

    def frames(self, n):
        self.substeps(n * self.pw.h("W_SUBSTEPS"))

    def ptrs(self):
        """The float oracle's buffers of this world as the C runners take them."""
        assert not self.f64
        pw = self.pw
        return WorldPtrs(*(P(a).value for a in (pw.header, pw.statF, self.statU, self.poses, self.vel, self.art, self.akey, self.aval,
                                                self.warm, self.count, self.stats)))

    def anchors_u32(self):
        assert not self.f64
        AH = self.pw.h("W_AH")
        out = np.zeros(2 * AH + 2 * AH * 14, np.uint32)
        lib().oracle_anchor_export_f32(ctypes.c_uint32(AH), P(self.akey), P(self.aval), P(out))
        return out

    def warm_u32(self):
        """The float oracle's contact warm-start table in the kernels' layout (keys, then 4 words per slot)."""
        assert not self.f64
        WH = self.pw.h("W_WH")
        out = np.zeros(10 * WH, np.uint32)
        lib().oracle_warm_export_f32(ctypes.c_uint32(WH), P(self.warm), P(out))
        return out

    def warm_map(self, parity=None):
        """{key: (normal, tangent world vector)} of one parity's table (default: the one the next substep reads)."""
        WH = self.pw.h("W_WH")
        if parity is None:
            parity = int(self.count[0]) % 2
        if self.f64:
            raise NotImplementedError
        return warm_table_map(self.warm_u32(), WH, parity)


def run_batch(worlds, frames_untimed, frames_timed, nthreads):
    """Step float-oracle worlds on nthreads std::threads (oracle_run_batch_f32): all worlds frames_untimed
    frames, then all worlds frames_timed frames; returns (wall seconds of the timed phase, busy seconds per
    thread)."""
    L = lib()
    arr = (WorldPtrs * len(worlds))(*[w.ptrs() for w in worlds])
    wall = ctypes.c_double()
    busy = (ctypes.c_double * max(nthreads, 1))()
    L.oracle_run_batch_f32(ctypes.c_uint32(len(worlds)), arr, ctypes.c_uint32(frames_untimed), ctypes.c_uint32(frames_timed),
                           ctypes.c_uint32(nthreads), ctypes.byref(wall), busy)
    return wall.value, list(busy)


def step_pair(o, nsub):
    """From float-oracle world o's current state, nsub substeps in float (o, in place) and in double (returned as
    (poses, vel, art) float64 arrays), plus the float and double wall times."""
    pw = o.pw
    p64 = np.zeros(pw.h("W_NPOSES") * 4)
    v64 = np.zeros(max(pw.NB, 1) * 8)
    a64 = np.zeros(pw.h("W_NART"))
    walls = np.zeros(2)
    ptr = o.ptrs()
    lib().oracle_step_pair_f32_f64(ctypes.byref(ptr), ctypes.c_uint32(nsub), P(p64), P(v64), P(a64), P(walls))
    return p64, v64, a64, walls


def state_quantities(pw, poses, vel, art):
    """Per-quantity views of a world's state: body COM positions and orientations, body twists (free bodies
    from vel, articulated bodies from their link twists), joint coordinates q and velocities qd."""
    from steel_webgpu_amd import pack as _pack
    NB = pw.NB
    P_ = np.asarray(poses, np.float64)[:NB * 8].reshape(NB, 8) if NB else np.zeros((0, 8))
    V = np.asarray(vel, np.float64)[:NB * 8].reshape(NB, 8).copy() if NB else np.zeros((0, 8))
    A = pw.A
    art = np.asarray(art, np.float64)
    for l, lk in enumerate(pw.links):
        o = A["A_LINK"] + l * _pack.LB + _pack.L_V
        b = lk["body"]
        V[b] = [art[o + 3], art[o + 4], art[o + 5], 0, art[o], art[o + 1], art[o + 2], 0]
    return {"position": P_[:, 0:3], "orientation": P_[:, 4:8], "linear_velocity": V[:, 0:3], "angular_velocity": V[:, 4:7],
            "q": art[A["A_Q"]:A["A_Q"] + pw.h("W_NQ")], "qd": art[A["A_QD"]:A["A_QD"] + pw.ND]}


def one_step_errors(o, nsub=None):
    """Error of one float step against the double step from the same state: per quantity, the largest absolute
    difference and that over the largest double magnitude (norm-wise relative error).  Steps o in place."""
    pw = o.pw
    nsub = nsub or pw.h("W_SUBSTEPS")
    p64, v64, a64, walls = step_pair(o, nsub)
    q32 = state_quantities(pw, o.poses, o.vel, o.art)
    q64 = state_quantities(pw, p64, v64, a64)
    out = {}
    for k in q32:
        if q64[k].size == 0:
            continue
        d = float(np.abs(q32[k] - q64[k]).max())
        m = float(np.abs(q64[k]).max())
        out[k] = {"abs": d, "rel": d / m if m > 0 else 0.0}
    return out, walls


def warm_table_map(buf, WH, parity):
    """{key: float32[4]} of one parity of a warm-start table in the kernels' uint32 layout."""
    keys = buf[parity * WH:(parity + 1) * WH]
    vals = buf[2 * WH:].reshape(2 * WH, 4)[parity * WH:(parity + 1) * WH].view(np.float32)
    return {int(k): vals[i].copy() for i, k in enumerate(keys) if k}


def art_probe(pw, mode, a0=0, a1=0, inp=None, nout=72, art=None, poses=None, vel=None):
    """The FP64 oracle's Featherstone functions on a world's state (initial state by default): FK and factor of
    every articulation, then mode 0 link response (W2, W), 1 dof_hinv, 2 art_solve_gen, 3 add_gen_jac, 4 link
    frames (see oracle_art_probe_f64)."""
    art = (pw.art0 if art is None else art).astype(np.float64).copy()
    poses = (pw.poses0 if poses is None else poses).astype(np.float64).copy()
    vel = (pw.vel0 if vel is None or not np.size(vel) else vel).astype(np.float64).copy()
    if not vel.size:
        vel = np.zeros(8)
    statU = pw.statU if pw.statU.size else np.zeros(1, np.uint32)
    out = np.zeros(nout)
    inp_arr = None if inp is None else np.ascontiguousarray(inp, np.float64)
    rc = lib().oracle_art_probe_f64(P(pw.header), P(pw.statF), P(statU), P(poses), P(vel), P(art), ctypes.c_uint32(mode),
                                    ctypes.c_uint32(a0), ctypes.c_uint32(a1), None if inp_arr is None else P(inp_arr), P(out))
    assert rc == 0
    return out


def mpr(pw, ha, hb, rq, t, r):
    """The oracle's MPR on hulls ha, hb of a packed world (hb rotated by rq, offset t in ha's frame)."""
    out = np.zeros(4)
    statU = pw.statU if pw.statU.size else np.zeros(1, np.uint32)
    lib().oracle_mpr_f64(P(pw.header), P(pw.statF), P(statU), ha, hb, P(np.asarray(rq, np.float64)), P(np.asarray(t, np.float64)), r, P(out))
    return out


def gsup(pw, shape, d):
    """The oracle's support point of a shape's collision geometry along d (shape frame)."""
    out = np.zeros(3)
    statU = pw.statU if pw.statU.size else np.zeros(1, np.uint32)
    lib().oracle_gsup_f64(P(pw.header), P(pw.statF), P(statU), shape, P(np.asarray(d, np.float64)), P(out))
    return out
