// Oracle: sequential CPU restatement of the reference's per-substep step
// (packages/steel-webgpu/src/wgsl.js and artwgsl.js), on the default option path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels: tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it; the product never links or calls it.
//
// It reads the packed world format of include/steel_world.h.  Every parallel phase of the reference is a
// plain loop here, in the order that phase's semantics fix (emission order is atomic-order in the
// reference and irrelevant after the canonical contact sort).  Templated on the scalar: float is the parity
// oracle (compiled with -ffp-contract=off, WGSL operation order), double is the error reference.
// Lane-strided SAT searches keep the reference's 32-lane tie order (wgsl.js:199-259, 1051-1061, 1114-1122).
#include "omath.h"
#include "../include/steel_world.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>
#include <vector>

namespace oracle {

static const uint32_t NLANE = 32;
static const uint32_t GT_PLANE = 1, GT_SPHERE = 3, GT_CAPSULE = 4, GT_CYL = 6, GT_BOX = 7, GT_CONE = 9, GT_CONVEX = 10;
static const uint32_t JT_PRISMATIC = 0, JT_REVOLUTE = 1, JT_BALL = 2, JT_FREE = 4, JT_D6 = 6;
// artwgsl.js:8-26 link block layout
static const uint32_t LB = 268, L_S = 0, L_U = 36, L_DI = 72, L_IA = 108, L_C = 135, L_PA = 141, L_V = 147, L_W = 153, L_A = 189,
                      L_UU = 195, L_P = 201, L_DV = 207, L_POS = 213, L_ROT = 216, L_COM = 220, L_TAU = 223, L_JR = 229, L_W2 = 232;
static const uint32_t AVN = 14;

template <class T> struct SV { V3<T> a, l; };
template <class T> inline SV<T> sv0() { return {{0, 0, 0}, {0, 0, 0}}; }
template <class T> inline SV<T> sadd(SV<T> x, SV<T> y) { return {x.a + y.a, x.l + y.l}; }
template <class T> inline SV<T> sscale(SV<T> x, T s) { return {x.a * s, x.l * s}; }
template <class T> inline T sdot(SV<T> m, SV<T> f) { return dot(m.a, f.a) + dot(m.l, f.l); }
template <class T> inline SV<T> crm(SV<T> v, SV<T> m) { return {cross(v.a, m.a), cross(v.a, m.l) + cross(v.l, m.a)}; }
template <class T> inline SV<T> crf(SV<T> v, SV<T> f) { return {cross(v.a, f.a) + cross(v.l, f.l), cross(v.a, f.l)}; }
template <class T> inline SV<T> mshift(SV<T> m, V3<T> r) { return {m.a, m.l + cross(m.a, r)}; }
template <class T> inline SV<T> fshift(SV<T> f, V3<T> r) { return {f.a + cross(r, f.l), f.l}; }
template <class T> inline M3<T> skew3(V3<T> r) { M3<T> m; m.c[0] = {0, r.z, -r.y}; m.c[1] = {-r.z, 0, r.x}; m.c[2] = {r.y, -r.x, 0}; return m; }
template <class T> inline M3<T> outer3(V3<T> x, V3<T> y) { M3<T> m; m.c[0] = x * y.x; m.c[1] = x * y.y; m.c[2] = x * y.z; return m; }
template <class T> inline M3<T> m3zero() { M3<T> m; m.c[0] = m.c[1] = m.c[2] = {0, 0, 0}; return m; }

// A u32 stored in a row float (wgsl.js:1723 bitcast<f32>(pair)): bits for float, value for double.
inline float put_u32(float*, uint32_t u) { return u2f(u); }
inline double put_u32(double*, uint32_t u) { return (double)u; }
inline uint32_t get_u32(float f) { return f2u(f); }
inline uint32_t get_u32(double d) { return (uint32_t)d; }
inline uint32_t fbits(float f) { return f2u(f); }
inline uint32_t fbits(double d) { return f2u((float)d); }
inline uint32_t fflip_u(uint32_t i) { return i ^ (((i >> 31) != 0 ? 0xffffffffu : 0u) | 0x80000000u); }

template <class T> struct Contact { V4<T> q[5]; T smu[2]; };
template <class T> struct AnchorVal { uint32_t pair, k; V3<T> ll, el, il, nl; };
// Contact warm-start table slot: row key, then the normal impulse and the tangent impulse as a world vector.
template <class T> struct WarmSlot { uint32_t key; V4<T> v; };
template <class T> struct FaceQ { T sep; uint32_t face; };
template <class T> struct EdgeQ { T sep; V3<T> axis, a0, a1, b0, b1; uint32_t id; };
template <class T> struct SegQ { T d2; V3<T> ps, ph; uint32_t face, kind; };
template <class T> struct Geo { uint32_t t; V3<T> s; uint32_t h; };
template <class T> struct Trk { V3<T> r, p, n; T a2; bool np; };
template <class T> struct Proj { T d; V3<T> n; };
template <class T> struct Man { V3<T> p[5]; T d[5]; uint32_t n; };

template <class T> struct World {
  // header
  const uint32_t* H;
  uint32_t NB, NS, NH, NA, NL, ND, NQ, NDA, JM, MAXPATH, MAXLP, MAXP, MAXC, MAXU, UH, AH, MAXDENSE, NCOLW, NCOL, ROWS, PO, PATCH, EXTRA,
      FCOMB, RCOMB, REST, ITERS, DHDR, DPS;
  uint32_t SF_BODY, SF_SHAPE, SF_HAABB, SF_VERTS, SF_FPLANE, SF_LINK, SF_DOF;
  uint32_t SU_HULL, SU_FACE, SU_FV, SU_HEDGE, SU_EDGE, SU_ART, SU_FILTER, FILT_LIST;
  uint32_t P_BODY, P_SHAPE, P_AABB, P_PREV;
  uint32_t A_Q, A_QD, A_QDH, A_TGT, A_TGTV, A_LIM, A_DRV, A_JF, A_BF, A_FRIC, A_HINV, A_LINK, A_LG, A_RL, A_RLN, A_ULOAD, A_DENSE, A_VLG, A_AVL;
  uint32_t WH, NEST, MO, TORS, TO, VELITERS, A_SNAP, COMPL, CO, DRIVE_ROWS, FUSE_VLIM, VEL_DRIVE_ACTIVE, NLOOP, SF_LOOP, ANG, CTRL_KIND, CTRL_A, CTRL_B, SF_CTRL, FMODE, FMO, ANG_ART, EO;
  T F_CREEP;
  bool vpass = false;  // velocity pass: rows without position bias or regularization
  T mom_k = 0, mom_beta = 0;  // Nesterov restart counter and extrapolation factor of the current substep
  uint32_t VLIM_ON, VLIM_FREE, JFL, A_JFL, SF_PATH, DRIVEN, PATH_FRAMES, INTEGRAL_COUNT, TGT_HELD, A_INT, SLEEP, A_SLP, NMIMIC, SF_MIMIC, SU_ACON;
  T FPS, FRAME_DT, F_SLEEP_LIN, F_SLEEP_ANG, F_SLEEP_TIME;
  T F_VLIM_FRAC;
  T F_CFM, F_OMEGA, F_SPEC, F_GAP_GATE, F_FGAP, F_FSCALE, F_ANCHOR_BETA, F_BP_CELL, F_REST_SPEED, F_JL_GAP, F_JL_SPEED;
  T dt, beta, max_depen, warm_scale, reg;
  V3<T> g;
  uint32_t parity;
  // Frame clock: substep index within the frame, control flags (bit 0: user-set targets), seconds since reset
  // (the reference's Sim.phase, accumulated in double as the host does).
  uint32_t substep_in_frame = 0, control = 0;
  double phase = 0;
  // buffers
  std::vector<V4<T>> statF;
  const uint32_t* statU;
  V4<T>* poses;
  V4<T>* vel;
  T* art;
  uint32_t* akey;           // [2*AH]
  AnchorVal<T>* aval;       // [2*AH]
  WarmSlot<T>* wtab;        // [2*WH] contact warm start (two parities)
  // scratch
  uint32_t counters[2];
  std::vector<uint32_t> pairs;  // 2 per pair
  std::vector<Contact<T>> craw, contacts;
  std::vector<V4<T>> rows;      // MAXC*ROWS + 2*MAXC slot table
  // sched sections (sim.js:442-444), one vector each
  // Body-pair keys (b + 1, a + 1) as one 64-bit word: the order of the reference's (a + 1) | (b + 1) << 16
  // (wgsl.js:1517-1524) for any body count.
  std::vector<uint64_t> S_HKEY, S_UKEY, wkey;
  std::vector<uint32_t> S_HCNT, S_SLOT2U, S_CSLOT, S_CLOCAL, S_USIZE_TMP, S_USIZE, S_USTART, S_CSTART, S_META, S_CORDER,
      S_RK0, S_RK1, S_RK2, S_KEPT, S_OSTART, S_LEAD, S_SLOT, S_ROWOF, S_RUNIT, S_PCOL, S_PORD, S_UTC, S_UTP, S_RCON, S_UWOFF, S_WALK;
  std::vector<uint64_t> wutp;   // coloring: each unit's largest (color + 1, walk index)
  std::vector<uint32_t> wcol, wmask, wcount, wlinkhit, wjsplit, uf, uff;
  uint32_t wnu, wnwalk, ndc;
  std::vector<V4<T>> tw, imp;
  std::vector<T> aqd;
  uint32_t stats[16];

  T hf(uint32_t k) const { return (T)u2f(H[k]); }
  void load_header(const uint32_t* h) {
    H = h;
    NB = h[W_NB]; NS = h[W_NS]; NH = h[W_NH]; NA = h[W_NA]; NL = h[W_NL]; ND = h[W_ND]; NQ = h[W_NQ]; NDA = h[W_NDA]; JM = h[W_JM];
    MAXPATH = h[W_MAXPATH]; MAXLP = h[W_MAXLP]; MAXP = h[W_MAXP]; MAXC = h[W_MAXC]; MAXU = h[W_MAXU]; UH = h[W_UH]; AH = h[W_AH];
    MAXDENSE = h[W_MAXDENSE]; NCOLW = h[W_NCOLW]; NCOL = NCOLW * 32; ROWS = h[W_ROWS]; PO = h[W_PO]; PATCH = h[W_PATCH]; EXTRA = h[W_EXTRA];
    FCOMB = h[W_FRICTION_COMBINE]; RCOMB = h[W_RESTITUTION_COMBINE]; REST = h[W_REST]; ITERS = h[W_ITERS]; DHDR = h[W_DHDR]; DPS = h[W_DPS];
    SF_BODY = h[W_SF_BODY]; SF_SHAPE = h[W_SF_SHAPE]; SF_HAABB = h[W_SF_HAABB]; SF_VERTS = h[W_SF_VERTS]; SF_FPLANE = h[W_SF_FPLANE];
    SF_LINK = h[W_SF_LINK]; SF_DOF = h[W_SF_DOF];
    SU_HULL = h[W_SU_HULL]; SU_FACE = h[W_SU_FACE]; SU_FV = h[W_SU_FV]; SU_HEDGE = h[W_SU_HEDGE]; SU_EDGE = h[W_SU_EDGE]; SU_ART = h[W_SU_ART];
    SU_FILTER = h[W_SU_FILTER];
    FILT_LIST = h[W_FILT_LIST];
    P_BODY = h[W_P_BODY]; P_SHAPE = h[W_P_SHAPE]; P_AABB = h[W_P_AABB]; P_PREV = h[W_P_PREV];
    A_Q = h[W_A_Q]; A_QD = h[W_A_QD]; A_QDH = h[W_A_QDH]; A_TGT = h[W_A_TGT]; A_TGTV = h[W_A_TGTV]; A_LIM = h[W_A_LIM]; A_DRV = h[W_A_DRV];
    A_JF = h[W_A_JF]; A_BF = h[W_A_BF]; A_FRIC = h[W_A_FRIC]; A_HINV = h[W_A_HINV]; A_LINK = h[W_A_LINK]; A_LG = h[W_A_LG]; A_RL = h[W_A_RL];
    A_RLN = h[W_A_RLN]; A_ULOAD = h[W_A_ULOAD]; A_DENSE = h[W_A_DENSE]; A_VLG = h[W_A_VLG]; A_AVL = h[W_A_AVL];
    WH = h[W_WH]; NEST = h[W_NEST]; MO = h[W_MO]; TORS = h[W_TORS]; TO = h[W_TO]; VELITERS = h[W_VELITERS]; A_SNAP = h[W_A_SNAP]; COMPL = h[W_COMPL]; CO = h[W_CO];
    NLOOP = h[W_NLOOP]; SF_LOOP = h[W_SF_LOOP]; ANG = h[W_ANG];
    FMODE = h[W_FMODE]; FMO = h[W_FMO]; ANG_ART = h[W_ANG_ART]; EO = h[W_EO]; F_CREEP = hf(W_F_CREEP);
    CTRL_KIND = h[W_CTRL_KIND]; CTRL_A = h[W_CTRL_A]; CTRL_B = h[W_CTRL_B]; SF_CTRL = h[W_SF_CTRL];
    DRIVE_ROWS = h[W_DRIVE_ROWS]; FUSE_VLIM = h[W_FUSE_VLIM]; VEL_DRIVE_ACTIVE = h[W_VEL_DRIVE_ACTIVE];
    VLIM_ON = h[W_VLIM_ON]; F_VLIM_FRAC = hf(W_F_VLIM_FRAC); VLIM_FREE = h[W_VLIM_FREE]; JFL = h[W_JFL]; A_JFL = h[W_A_JFL];
    SF_PATH = h[W_SF_PATH]; DRIVEN = h[W_DRIVEN]; PATH_FRAMES = h[W_PATH_FRAMES]; INTEGRAL_COUNT = h[W_INTEGRAL_COUNT];
    TGT_HELD = h[W_TGT_HELD]; A_INT = h[W_A_INT]; FPS = hf(W_FPS); FRAME_DT = hf(W_FRAME_DT);
    SLEEP = h[W_SLEEP]; A_SLP = h[W_A_SLP]; NMIMIC = h[W_NMIMIC]; SF_MIMIC = h[W_SF_MIMIC]; SU_ACON = h[W_SU_ACON]; F_SLEEP_LIN = hf(W_F_SLEEP_LIN); F_SLEEP_ANG = hf(W_F_SLEEP_ANG); F_SLEEP_TIME = hf(W_F_SLEEP_TIME);
    F_CFM = hf(W_F_CFM); F_OMEGA = hf(W_F_OMEGA); F_SPEC = hf(W_F_SPEC); F_GAP_GATE = hf(W_F_GAP_GATE); F_FGAP = hf(W_F_FGAP);
    F_FSCALE = hf(W_F_FSCALE); F_ANCHOR_BETA = hf(W_F_ANCHOR_BETA); F_BP_CELL = hf(W_F_BP_CELL); F_REST_SPEED = hf(W_F_REST_SPEED);
    F_JL_GAP = hf(W_F_JL_GAP); F_JL_SPEED = hf(W_F_JL_SPEED);
    dt = hf(W_DT); beta = hf(W_BETA); max_depen = hf(W_MAX_DEPEN); warm_scale = hf(W_WARM_SCALE); reg = hf(W_REG);
    g = {hf(W_GX), hf(W_GY), hf(W_GZ)};
  }

  // ---- statF/statU accessors (wgsl.js:70-105, 185-195; artwgsl.js:31-55)
  V4<T> sf(uint32_t i) const { return statF[i]; }
  T body_inv_mass(int b) const { return statF[SF_BODY + (uint32_t)b * 8].x; }
  uint32_t body_kind(int b) const { return tu(statF[SF_BODY + (uint32_t)b * 8].y); }
  int body_art(int b) const { return ti(statF[SF_BODY + (uint32_t)b * 8].z); }
  int body_link(int b) const { return ti(statF[SF_BODY + (uint32_t)b * 8].w); }
  V3<T> body_com_local(int b) const { return xyz(statF[SF_BODY + (uint32_t)b * 8 + 1]); }
  bool same_art(int a, int b) const { return a >= 0 && b >= 0 && body_kind(a) == 2 && body_kind(b) == 2 && body_art(a) == body_art(b); }
  T body_mass(int b) const { return statF[SF_BODY + (uint32_t)b * 8 + 1].w; }
  V4<T> body_limits(int b) const {
    uint32_t o = SF_BODY + (uint32_t)b * 8;
    return {statF[o + 3].w, statF[o + 4].w, statF[o + 5].w, statF[o + 6].w};
  }
  T depen_cap(int a, int b) const {
    T m = T(1e20);
    if (a >= 0) { T l = body_limits(a).w; if (l > T(0)) m = mn(m, l); }
    if (b >= 0) { T l = body_limits(b).w; if (l > T(0)) m = mn(m, l); }
    return m;
  }
  uint32_t body_flags(int b) const { if (b < 0) return 0; return tu(statF[SF_BODY + (uint32_t)b * 8 + 7].w); }
  T body_radius(int b) const { if (b < 0) return T(1e30); return statF[SF_BODY + (uint32_t)b * 8 + 2].w; }
  M3<T> body_rows_T(uint32_t o) const {  // transpose(mat3x3f(row0, row1, row2))
    M3<T> m; m.c[0] = xyz(statF[o]); m.c[1] = xyz(statF[o + 1]); m.c[2] = xyz(statF[o + 2]);
    return transpose(m);
  }
  M3<T> body_inertia(int b) const { return body_rows_T(SF_BODY + (uint32_t)b * 8 + 2); }
  M3<T> body_inv_inertia(int b) const { return body_rows_T(SF_BODY + (uint32_t)b * 8 + 5); }
  V3<T> hv(uint32_t h, uint32_t i) const { return xyz(statF[SF_VERTS + statU[SU_HULL + h * 4] + i]); }
  uint32_t hnv(uint32_t h) const { return statU[SU_HULL + h * 4 + 1]; }
  uint32_t hfo(uint32_t h) const { return statU[SU_HULL + h * 4 + 2]; }
  uint32_t hnf(uint32_t h) const { return statU[SU_HULL + h * 4 + 3]; }
  V4<T> fplane(uint32_t f) const { return statF[SF_FPLANE + f]; }
  uint32_t fstart(uint32_t f) const { return statU[SU_FACE + f * 2]; }
  uint32_t fcount(uint32_t f) const { return statU[SU_FACE + f * 2 + 1]; }
  uint32_t fvert(uint32_t k) const { return statU[SU_FV + k]; }
  uint32_t hne(uint32_t h) const { return statU[SU_HEDGE + h * 2 + 1]; }
  void hedge(uint32_t h, uint32_t k, uint32_t e[4]) const {
    uint32_t o = SU_EDGE + (statU[SU_HEDGE + h * 2] + k) * 4;
    e[0] = statU[o]; e[1] = statU[o + 1]; e[2] = statU[o + 2]; e[3] = statU[o + 3];
  }
  V4<T> shape_prim(uint32_t s) const { return statF[SF_SHAPE + s * 7 + 3]; }
  uint32_t geo_type(uint32_t s) const { return tu(statF[SF_SHAPE + s * 7 + 6].x); }
  Geo<T> geo(uint32_t s) const {
    V4<T> gg = statF[SF_SHAPE + s * 7 + 6];
    Geo<T> r{tu(gg.x), {gg.y, gg.z, gg.w}, tu(statF[SF_SHAPE + s * 7].y)};
    if (r.t == GT_SPHERE || r.t == GT_CAPSULE) r.s.x = T(1e-4);
    return r;
  }
  // articulation records
  uint32_t link_type(uint32_t l) const { return tu(statF[SF_LINK + l * 6].x); }
  int link_parent(uint32_t l) const { return ti(statF[SF_LINK + l * 6].y); }
  uint32_t link_dof(uint32_t l) const { return tu(statF[SF_LINK + l * 6].z); }
  int link_body(uint32_t l) const { return ti(statF[SF_LINK + l * 6].w); }
  uint32_t link_ndof(uint32_t l) const { return tu(statF[SF_LINK + l * 6 + 1].w); }
  uint32_t link_qs(uint32_t l) const { return tu(statF[SF_LINK + l * 6 + 3].w); }
  V4<T> dof_drive(uint32_t d) const { return statF[SF_DOF + d * 5]; }
  bool driven_dof(uint32_t d) const { V4<T> dr = dof_drive(d); return dr.y > T(0) || dr.z > T(0); }
  // drive_mode rows: a driven dof's velocity limit is left to the sweep clamp (artwgsl.js:54)
  bool fused_dof(uint32_t d) const { return DRIVE_ROWS && FUSE_VLIM && driven_dof(d); }
  V3<T> dof_axis(uint32_t d) const { return xyz(statF[SF_DOF + d * 5 + 3]); }
  uint32_t link_nlin(uint32_t l) const { return tu(statF[SF_LINK + l * 6 + 5].w); }  // D6: leading translation dofs
  V4<T> dof_limit(uint32_t d) const { return statF[SF_DOF + d * 5 + 1]; }
  V4<T> dof_passive(uint32_t d) const { return statF[SF_DOF + d * 5 + 2]; }
  bool scalar_dof(uint32_t d) const { return statF[SF_DOF + d * 5 + 2].w > T(0.5); }
  void art_info(uint32_t a, uint32_t ai[4]) const { for (int k = 0; k < 4; k++) ai[k] = statU[SU_ART + a * 4 + k]; }
  uint32_t lo(uint32_t l) const { return A_LINK + l * LB; }
  V3<T> ld3(uint32_t o) const { return {art[o], art[o + 1], art[o + 2]}; }
  void st3(uint32_t o, V3<T> v) { art[o] = v.x; art[o + 1] = v.y; art[o + 2] = v.z; }
  SV<T> ldsv(uint32_t o) const { return {ld3(o), ld3(o + 3)}; }
  void stsv(uint32_t o, SV<T> s) { st3(o, s.a); st3(o + 3, s.l); }
  M3<T> ldm(uint32_t o) const { M3<T> m; m.c[0] = ld3(o); m.c[1] = ld3(o + 3); m.c[2] = ld3(o + 6); return m; }
  void stm(uint32_t o, const M3<T>& m) { st3(o, m.c[0]); st3(o + 3, m.c[1]); st3(o + 6, m.c[2]); }
  SV<T> imul(uint32_t o, SV<T> m) const {
    M3<T> A = ldm(o), B = ldm(o + 9), C = ldm(o + 18);
    return {mul(A, m.a) + mul(B, m.l), mul(transpose(B), m.a) + mul(C, m.l)};
  }
  SV<T> ldS(uint32_t l, uint32_t k) const { return ldsv(lo(l) + L_S + k * 6); }
  SV<T> ldU(uint32_t l, uint32_t k) const { return ldsv(lo(l) + L_U + k * 6); }
  T ldDI(uint32_t l, uint32_t r, uint32_t c) const { return art[lo(l) + L_DI + r * JM + c]; }
  V3<T> link_r(uint32_t l) const {
    int par = link_parent(l);
    if (par < 0) return {0, 0, 0};
    return ld3(lo(l) + L_COM) - ld3(lo((uint32_t)par) + L_COM);
  }
  V4<T> link_rot(uint32_t l) const { V3<T> r = ld3(lo(l) + L_ROT); return {r.x, r.y, r.z, art[lo(l) + L_ROT + 3]}; }

  // ============================ shapes (wgsl.js:117-150) ============================
  void shapes() {
    for (uint32_t s = 0; s < NS; s++) {
      uint32_t o = SF_SHAPE + s * 7;
      int body = ti(statF[o].x);
      uint32_t hull = tu(statF[o].y);
      V3<T> bp{0, 0, 0};
      V4<T> bq{0, 0, 0, 1};
      if (body >= 0) {
        V3<T> c = xyz(poses[P_BODY + (uint32_t)body * 2]);
        bq = poses[P_BODY + (uint32_t)body * 2 + 1];
        bp = c - qrot(bq, body_com_local(body));
      }
      V3<T> sp = bp + qrot(bq, xyz(statF[o + 1]));
      V4<T> sq = normalize4(qmul(bq, statF[o + 2]));
      poses[P_SHAPE + s * 2] = v4(sp, T(0));
      poses[P_SHAPE + s * 2 + 1] = sq;
      if (statF[o + 3].x == T(4)) {
        poses[P_AABB + s * 2] = {T(-1e6), T(-1e6), T(-1e6), 0};
        poses[P_AABB + s * 2 + 1] = {T(1e6), T(1e6), T(1e6), 0};
        continue;
      }
      V3<T> hc = xyz(statF[SF_HAABB + hull * 2]);
      V3<T> hh = xyz(statF[SF_HAABB + hull * 2 + 1]);
      M3<T> m = qmat(sq);
      V3<T> c = sp + mul(m, hc);
      V3<T> e = vabs(m.c[0]) * hh.x + vabs(m.c[1]) * hh.y + vabs(m.c[2]) * hh.z + V3<T>{statF[o].w, statF[o].w, statF[o].w};
      poses[P_AABB + s * 2] = v4(c - e, T(0));
      poses[P_AABB + s * 2 + 1] = v4(c + e, T(0));
    }
  }

  // ============================ broadphase (wgsl.js:154-177) ============================
  // Worlds of more than 4096 shapes find the same pair set by a sweep over the shapes sorted by AABB min x (the
  // candidates of shape i are the later shapes whose min x is at most i's max x); only the emission order differs.
  void broad() {
    pairs.clear();
    counters[0] = 0;
    if (NS > 4096) {
      std::vector<uint32_t> ord(NS);
      for (uint32_t i = 0; i < NS; i++) ord[i] = i;
      std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return poses[P_AABB + x * 2].x < poses[P_AABB + y * 2].x; });
      for (uint32_t u = 0; u < NS; u++)
        for (uint32_t v = u + 1; v < NS; v++) {
          if (poses[P_AABB + ord[v] * 2].x > poses[P_AABB + ord[u] * 2 + 1].x) break;
          broad_test(std::min(ord[u], ord[v]), std::max(ord[u], ord[v]));
        }
      return;
    }
    for (uint32_t i = 0; i < NS; i++)
      for (uint32_t j = i + 1; j < NS; j++) broad_test(i, j);
  }
  void broad_test(uint32_t i, uint32_t j) {
      {
        int bi = ti(statF[SF_SHAPE + i * 7].x), bj = ti(statF[SF_SHAPE + j * 7].x);
        if (bi == bj) return;
        if (bi < 0 && bj < 0) return;
        bool fi = bi < 0 || body_kind(bi) == 3 || (body_flags(bi) & 1) != 0;
        bool fj = bj < 0 || body_kind(bj) == 3 || (body_flags(bj) & 1) != 0;
        if (fi && fj) return;
        if (FILT_LIST) {  // shape i's sorted list of filtered partners j > i
          const uint32_t* L = statU + SU_FILTER + NS + 1;
          if (std::binary_search(L + statU[SU_FILTER + i], L + statU[SU_FILTER + i + 1], j)) return;
        } else {
          uint64_t k = (uint64_t)i * NS + j;
          if ((statU[SU_FILTER + (uint32_t)(k >> 5)] & (1u << (uint32_t)(k & 31))) != 0) return;
        }
        V3<T> lo_i = xyz(poses[P_AABB + i * 2]), hi_i = xyz(poses[P_AABB + i * 2 + 1]);
        V3<T> lo_j = xyz(poses[P_AABB + j * 2]), hi_j = xyz(poses[P_AABB + j * 2 + 1]);
        if (lo_i.x > hi_j.x || lo_i.y > hi_j.y || lo_i.z > hi_j.z || lo_j.x > hi_i.x || lo_j.y > hi_i.y || lo_j.z > hi_i.z) return;
        uint32_t slot = counters[0]++;
        if (slot < MAXP) { pairs.push_back(i); pairs.push_back(j); }
      }
  }

  // ============================ narrowphase (wgsl.js:182-1215) ============================
  FaceQ<T> face_query(uint32_t ha, uint32_t hb, V4<T> rq, V3<T> t, uint32_t lane) const {
    FaceQ<T> best{T(-1e30), 0};
    uint32_t fo = hfo(ha), nf = hnf(ha), nv = hnv(hb);
    for (uint32_t f = lane; f < nf; f += NLANE) {
      V4<T> pl = fplane(fo + f);
      V3<T> nb = qinvrot(rq, xyz(pl));
      T m = T(1e30);
      for (uint32_t v = 0; v < nv; v++) m = mn(m, dot(nb, hv(hb, v)));
      T s = m + dot(xyz(pl), t) - pl.w;
      if (s > best.sep) best = {s, f};
    }
    return best;
  }
  // reduce_face (wgsl.js:1051-1061): max sep, lowest lane on ties.
  template <class F> FaceQ<T> reduce_lanes_max(F fn) const {
    FaceQ<T> r[NLANE];
    for (uint32_t l = 0; l < NLANE; l++) r[l] = fn(l);
    uint32_t b = 0;
    for (uint32_t k = 1; k < NLANE; k++) if (r[k].sep > r[b].sep) b = k;
    return r[b];
  }
  T support_min(uint32_t h, V3<T> n) const { T m = T(1e30); for (uint32_t v = 0; v < hnv(h); v++) m = mn(m, dot(n, hv(h, v))); return m; }
  T support_max(uint32_t h, V3<T> n) const { T m = T(-1e30); for (uint32_t v = 0; v < hnv(h); v++) m = mx(m, dot(n, hv(h, v))); return m; }
  EdgeQ<T> edge_query(uint32_t ha, uint32_t hb, V4<T> rq, V3<T> t, uint32_t lane) const {
    V3<T> z{0, 0, 0};
    EdgeQ<T> best{T(-1e30), z, z, z, z, z, 0};
    uint32_t foa = hfo(ha), fob = hfo(hb);
    V3<T> cena = xyz(statF[SF_HAABB + ha * 2]);
    V3<T> cenb = qrot(rq, xyz(statF[SF_HAABB + hb * 2])) + t;
    uint32_t nea = hne(ha), neb = hne(hb);
    for (uint32_t k = lane; k < nea * neb; k += NLANE) {
      uint32_t ea_[4], eb_[4];
      hedge(ha, k / neb, ea_); hedge(hb, k % neb, eb_);
      V3<T> a = xyz(fplane(foa + ea_[2])), b = xyz(fplane(foa + ea_[3]));
      V3<T> c = -qrot(rq, xyz(fplane(fob + eb_[2]))), d = -qrot(rq, xyz(fplane(fob + eb_[3])));
      V3<T> bxa = cross(b, a), dxc = cross(d, c);
      T cba = dot(c, bxa), dba = dot(d, bxa), adc = dot(a, dxc), bdc = dot(b, dxc);
      if (cba * dba > T(1e-3) || adc * bdc > T(1e-3) || cba * bdc < T(-1e-3)) continue;
      uint32_t va0 = ea_[0], va1 = ea_[1], vb0 = eb_[0], vb1 = eb_[1];
      V3<T> a0 = hv(ha, va0), a1 = hv(ha, va1);
      V3<T> ea = a1 - a0;
      V3<T> b0 = qrot(rq, hv(hb, vb0)) + t, b1 = qrot(rq, hv(hb, vb1)) + t;
      V3<T> ax = cross(ea, b1 - b0);
      T l = length(ax);
      if (l < T(1e-6) * length(ea) * length(b1 - b0)) continue;
      ax = ax / l;
      if (dot(ax, cenb - cena) < T(0)) ax = -ax;
      T amax = support_max(ha, ax);
      T bmin = dot(ax, t) + support_min(hb, qinvrot(rq, ax));
      T s = bmin - amax;
      T tol = T(1e-4) + T(1e-3) * (length(ea) + length(b1 - b0));
      if (dot(ax, a0) < amax - tol || dot(ax, a1) < amax - tol || dot(ax, b0) > bmin + tol || dot(ax, b1) > bmin + tol) continue;
      if (s > best.sep) best = {s, ax, a0, a1, b0, b1, (va0 << 24) ^ (va1 << 16) ^ (vb0 << 8) ^ vb1};
    }
    return best;
  }
  static T combine(T a, T b, uint32_t mode) {
    if (mode == 1) return mn(a, b);
    if (mode == 2) return a * b;
    if (mode == 3) return mx(a, b);
    return T(0.5) * (a + b);
  }
  // wgsl.js:269-307
  void emit(uint32_t sa, uint32_t sb, V3<T> p, V3<T> n, T phi, uint32_t key) {
    int ba = ti(statF[SF_SHAPE + sa * 7].x), bb = ti(statF[SF_SHAPE + sb * 7].x);
    V4<T> ma = statF[SF_SHAPE + sa * 7 + 4], mb = statF[SF_SHAPE + sb * 7 + 4];
    T mu = combine(statF[SF_SHAPE + sa * 7].z, statF[SF_SHAPE + sb * 7].z, FCOMB);
    bool art_c = (ba >= 0 && body_kind(ba) >= 2) || (bb >= 0 && body_kind(bb) >= 2);
    T mt = 0, mr = 0;
    if (art_c) {
      if (ANG_ART) { mt = combine(ma.y, mb.y, FCOMB); mr = combine(ma.z, mb.z, FCOMB); }
    } else {
      mt = mx(ma.w >= T(4) ? ma.y : T(0), mb.w >= T(4) ? mb.y : T(0));
      mr = mx(ma.w >= T(6) ? ma.z : T(0), mb.w >= T(6) ? mb.z : T(0));
    }
    T tr = art_c ? mx(statF[SF_SHAPE + sa * 7 + 3].w, statF[SF_SHAPE + sb * 7 + 3].w) : T(0);
    V3<T> mat{combine(clampv(ma.x, T(0), T(1)), clampv(mb.x, T(0), T(1)), RCOMB), mt, mr};
    int a = ba, b = bb;
    V3<T> nn = n;
    T smu0 = statF[SF_SHAPE + sa * 7].z, smu1 = statF[SF_SHAPE + sb * 7].z;
    if (a > b) { a = bb; b = ba; nn = -n; std::swap(smu0, smu1); }
    uint32_t k = counters[1]++;
    if (k >= MAXC) return;
    Contact<T> c;
    c.smu[0] = smu0; c.smu[1] = smu1;
    c.q[0] = v4(p, phi);
    c.q[1] = v4(nn, mu);
    c.q[2] = {(T)a, (T)b, (T)(key & 0xffffu), (T)(key >> 16)};
    c.q[3] = v4(mat, tr);
    T ka = statF[SF_SHAPE + sa * 7 + 5].x, kb = statF[SF_SHAPE + sb * 7 + 5].x;
    T kc = (ka > T(0) && kb > T(0)) ? ka * kb / (ka + kb) : T(0);
    T ga = statF[SF_SHAPE + sa * 7 + 5].z, gb = statF[SF_SHAPE + sb * 7 + 5].z;
    if (ba > bb) std::swap(ga, gb);
    c.q[4] = {kc, statF[SF_SHAPE + sa * 7 + 5].y + statF[SF_SHAPE + sb * 7 + 5].y, ga, gb};
    craw.push_back(c);
  }
  static void closest_seg(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1, V3<T> out[2]) {  // wgsl.js:393-411
    V3<T> d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
    T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    T s = 0, t = 0;
    if (a <= T(1e-12) && e <= T(1e-12)) { out[0] = a0; out[1] = b0; return; }
    if (a <= T(1e-12)) t = clampv(f / e, T(0), T(1));
    else {
      T c = dot(d1, r);
      if (e <= T(1e-12)) s = clampv(-c / a, T(0), T(1));
      else {
        T b = dot(d1, d2), den = a * e - b * b;
        if (den > T(1e-12)) s = clampv((b * f - c * e) / den, T(0), T(1));
        t = (b * s + f) / e;
        if (t < T(0)) { t = 0; s = clampv(-c / a, T(0), T(1)); }
        else if (t > T(1)) { t = 1; s = clampv((b - c) / a, T(0), T(1)); }
      }
    }
    out[0] = a0 + d1 * s; out[1] = b0 + d2 * t;
  }
  static void sphere_sphere(V3<T> p1, T r1, V3<T> p2, T r2, V4<T> out[2]) {  // wgsl.js:414-420
    V3<T> dv = p2 - p1;
    T l = length(dv);
    V3<T> n = (l == T(0)) ? V3<T>{1, 0, 0} : dv / l;
    T d = l - (r1 + r2);
    out[0] = v4(p1 + n * (r1 + T(0.5) * d), d);
    out[1] = v4(n, T(0));
  }
  void prim_prim(uint32_t sa, uint32_t sb, V3<T> pa, V4<T> qa, V3<T> pb, V4<T> qb, T thresh) {  // wgsl.js:424-454
    V4<T> A = statF[SF_SHAPE + sa * 7 + 6], B = statF[SF_SHAPE + sb * 7 + 6];
    V3<T> ua = qrot(qa, V3<T>{0, 0, A.z}), ub = qrot(qb, V3<T>{0, 0, B.z});
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    V4<T> c[2][2];
    uint32_t nc = 1;
    if (tu(A.x) == GT_SPHERE) {
      V3<T> a = pb - ub, ab = (pb + ub) - a;
      sphere_sphere(pa, A.y, a + clampv(dot(pa - a, ab) / (dot(ab, ab) + T(1e-6)), T(0), T(1)) * ab, B.y, c[0]);
    } else {
      V3<T> dif = pa - pb;
      T ma = dot(ua, ua), mb = -dot(ua, ub), mc = dot(ub, ub);
      T u = -dot(ua, dif), v = dot(ub, dif);
      T det = ma * mc - mb * mb;
      if (std::fabs(det) >= T(1e-15)) {
        T inv_det = T(1) / det;
        T x1 = (mc * u - mb * v) * inv_det, x2 = (ma * v - mb * u) * inv_det;
        if (x1 > T(1)) { x1 = 1; x2 = (v - mb) / mc; } else if (x1 < T(-1)) { x1 = -1; x2 = (v + mb) / mc; }
        if (x2 > T(1)) { x2 = 1; x1 = clampv((u - mb) / ma, T(-1), T(1)); } else if (x2 < T(-1)) { x2 = -1; x1 = clampv((u + mb) / ma, T(-1), T(1)); }
        sphere_sphere(pa + ua * x1, A.y, pb + ub * x2, B.y, c[0]);
      } else {
        sphere_sphere(pa + ua, A.y, pb + ub * clampv((v - mb) / mc, T(-1), T(1)), B.y, c[0]);
        sphere_sphere(pa - ua, A.y, pb + ub * clampv((v + mb) / mc, T(-1), T(1)), B.y, c[1]);
        c[1][1] = c[0][1];
        nc = 2;
      }
    }
    for (uint32_t i = 0; i < nc; i++)
      if (c[i][0].w <= thresh) emit(sa, sb, xyz(c[i][0]), xyz(c[0][1]), c[i][0].w, hash_u32(base ^ (i * 977u + 13u)));
  }
  void plane_contacts(uint32_t pl, uint32_t os, V3<T> pp, V4<T> pq, V3<T> op, V4<T> oq, T thresh, bool flip) {  // wgsl.js:458-540
    V3<T> n = qrot(pq, V3<T>{0, 0, 1});
    uint32_t sa = flip ? os : pl, sb = flip ? pl : os;
    V3<T> nab = flip ? -n : n;
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    V4<T> g = statF[SF_SHAPE + os * 7 + 6];
    uint32_t gt = tu(g.x);
    T d[4] = {T(1e30), T(1e30), T(1e30), T(1e30)};
    V3<T> x[4];
    uint32_t nc = 0;
    if (gt == GT_SPHERE || gt == GT_CAPSULE) {
      V3<T> u = qrot(oq, V3<T>{0, 0, g.z});
      uint32_t cnt = gt == GT_SPHERE ? 1 : 2;
      for (uint32_t i = 0; i < cnt; i++) {
        V3<T> e = op + (i == 1 ? -u : u);
        d[i] = dot(n, e - pp) - g.y;
        x[i] = e - n * (g.y + T(0.5) * d[i]);
      }
      nc = cnt;
    } else if (gt == GT_BOX) {
      T cd0 = dot(op - pp, n);
      uint32_t worst = 0;
      M3<T> rot = qmat(oq);
      for (uint32_t i = 0; i < 8; i++) {
        V3<T> c = mul(rot, V3<T>{(i & 1) ? g.y : -g.y, (i & 2) ? g.z : -g.z, (i & 4) ? g.w : -g.w});
        T cd = cd0 + dot(n, c);
        if (cd > thresh) continue;
        V3<T> cp = c + op - (T(0.5) * n) * cd;
        if (nc < 4) {
          d[nc] = cd; x[nc] = cp;
          if (nc == 0 || cd > d[worst]) worst = nc;
          nc++;
        } else if (cd < d[worst]) {
          d[worst] = cd; x[worst] = cp;
          worst = 0;
          for (uint32_t k = 1; k < 4; k++) if (d[k] > d[worst]) worst = k;
        }
      }
    } else if (gt == GT_CYL) {
      // Deepest rim point, then three cap points (within 22.5 degrees of upright) or the opposite rim and the
      // cap's side point (rolling).
      V3<T> ax = qrot(oq, V3<T>{0, 0, 1});
      T dna = dot(n, ax);
      if (dna > T(0)) { ax = -ax; dna = -dna; }
      V3<T> cap = op + ax * g.z;
      V3<T> pa = -n + ax * dna;
      T l2 = dot(pa, pa);
      bool has = l2 > T(1e-10);
      if (has) pa = pa * (T(1) / std::sqrt(l2));
      bool flat = -dna >= T(0.9238795325112867);
      V3<T> pf{0, 0, 0};
      if (flat || !has) {
        V3<T> rf{1, 0, 0};
        if (std::fabs(dot(ax, rf)) > T(0.9)) rf = {0, 1, 0};
        pf = normalize(rf - ax * dot(ax, rf));
      }
      V3<T> dp = has ? pa : pf;
      V3<T> dpt = cap + dp * g.y;
      d[0] = dot(dpt - pp, n);
      x[0] = dpt - n * (d[0] * T(0.5));
      nc = 1;
      T mt = T(0.01) * mx(g.y, g.z);
      V3<T> cand[3];
      uint32_t ncand = 0;
      if (flat) {
        V3<T> u = pf * g.y, v = cross(ax, pf) * g.y;
        cand[0] = cap + u;
        cand[1] = cap - T(0.5) * u + T(0.8660254) * v;
        cand[2] = cap - T(0.5) * u - T(0.8660254) * v;
        ncand = 3;
      } else {
        V3<T> u = dp * g.y, v = cross(ax, dp) * g.y;
        cand[0] = op - ax * g.z + u;
        T dpv = dot(cap + v - pp, n), dnv = dot(cap - v - pp, n);
        cand[1] = dpv <= dnv ? cap + v : cap - v;
        ncand = 2;
      }
      for (uint32_t k = 0; k < ncand; k++) {
        T dk = dot(cand[k] - pp, n);
        V3<T> pk = cand[k] - n * (dk * T(0.5));
        if (nc < 4 && dot(pk - x[0], pk - x[0]) > mt * mt) { d[nc] = dk; x[nc] = pk; nc++; }
      }
    }
    for (uint32_t i = 0; i < nc; i++)
      if (d[i] <= thresh) emit(sa, sb, x[i], nab, d[i], hash_u32(base ^ (i + 7u)));
  }
  FaceQ<T> face_query_seg(uint32_t h, V3<T> e0, V3<T> e1, T r, uint32_t lane) const {  // wgsl.js:543-552
    FaceQ<T> best{T(-1e30), 0};
    uint32_t fo = hfo(h);
    for (uint32_t f = lane; f < hnf(h); f += NLANE) {
      V4<T> pl = fplane(fo + f);
      T s = mn(dot(xyz(pl), e0), dot(xyz(pl), e1)) - pl.w - r;
      if (s > best.sep) best = {s, f};
    }
    return best;
  }
  bool seg_hits_hull(uint32_t h, V3<T> e0, V3<T> e1) const {  // wgsl.js:554-563
    T t0 = 0, t1 = 1;
    for (uint32_t f = 0; f < hnf(h); f++) {
      V4<T> pl = fplane(hfo(h) + f);
      T d0 = dot(xyz(pl), e0) - pl.w, d1 = dot(xyz(pl), e1) - pl.w;
      if (d0 > T(0) && d1 > T(0)) return false;
      if (d0 > T(0)) t0 = mx(t0, d0 / (d0 - d1));
      else if (d1 > T(0)) t1 = mn(t1, d0 / (d0 - d1));
    }
    return t0 <= t1;
  }
  SegQ<T> seg_hull_query(uint32_t h, V3<T> e0, V3<T> e1, uint32_t lane) const {  // wgsl.js:568-590
    V3<T> z{0, 0, 0};
    SegQ<T> best{T(1e30), z, z, 0, 1};
    uint32_t fo = hfo(h);
    for (uint32_t f = lane; f < hnf(h); f += NLANE) {
      V4<T> pl = fplane(fo + f);
      uint32_t sr = fstart(fo + f), cr = fcount(fo + f);
      bool in0 = dot(xyz(pl), e0) > pl.w, in1 = dot(xyz(pl), e1) > pl.w;
      for (uint32_t k = 0; k < cr; k++) {
        V3<T> r0 = hv(h, fvert(sr + k)), r1 = hv(h, fvert(sr + (k + 1) % cr));
        V3<T> out = cross(r1 - r0, xyz(pl));
        if (dot(e0 - r0, out) > T(0)) in0 = false;
        if (dot(e1 - r0, out) > T(0)) in1 = false;
        V3<T> c[2];
        closest_seg(e0, e1, r0, r1, c);
        T d2 = dot(c[0] - c[1], c[0] - c[1]);
        if (d2 < best.d2) best = {d2, c[0], c[1], f, 1};
      }
      if (in0) { T sd = dot(xyz(pl), e0) - pl.w; if (sd * sd <= best.d2 * T(1.0001)) best = {sd * sd, e0, e0 - xyz(pl) * sd, f, 0}; }
      if (in1) { T sd = dot(xyz(pl), e1) - pl.w; if (sd * sd <= best.d2 * T(1.0001)) best = {sd * sd, e1, e1 - xyz(pl) * sd, f, 0}; }
    }
    return best;
  }
  void seg_on_face(uint32_t hs, uint32_t ps, uint32_t h, uint32_t f, V3<T> e0, V3<T> e1, T r, V3<T> wp, V4<T> wq, T thresh, bool flip) {
    V4<T> pl = fplane(hfo(h) + f);  // wgsl.js:594-624
    T t0 = 0, t1 = 1;
    uint32_t sr = fstart(hfo(h) + f), cr = fcount(hfo(h) + f);
    for (uint32_t k = 0; k < cr; k++) {
      V3<T> r0 = hv(h, fvert(sr + k)), r1 = hv(h, fvert(sr + (k + 1) % cr));
      V3<T> out = cross(r1 - r0, xyz(pl));
      T d0 = dot(e0 - r0, out), d1 = dot(e1 - r0, out);
      if (d0 > T(0) && d1 > T(0)) { t0 = 2; break; }
      if (d0 > T(0)) t0 = mx(t0, d0 / (d0 - d1));
      if (d1 > T(0)) t1 = mn(t1, d0 / (d0 - d1));
    }
    T ts[2] = {t0, t1};
    uint32_t n = 2;
    if (t0 > t1) { ts[0] = dot(xyz(pl), e0) <= dot(xyz(pl), e1) ? T(0) : T(1); n = 1; }
    else if (t1 - t0 < T(1e-4) || length(e1 - e0) < T(1e-7)) n = 1;
    V3<T> nw = qrot(wq, xyz(pl));
    uint32_t sa = flip ? ps : hs, sb = flip ? hs : ps;
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    for (uint32_t i = 0; i < n; i++) {
      V3<T> x = e0 + (e1 - e0) * ts[i];
      T sd = dot(xyz(pl), x) - pl.w;
      T phi = sd - r;
      if (phi > thresh) continue;
      V3<T> mid = x - xyz(pl) * (T(0.5) * (sd + r));
      emit(sa, sb, wp + qrot(wq, mid), flip ? -nw : nw, phi, hash_u32(base ^ (f * 131u + i * 7u + 3u)));
    }
  }
  // ---- generic manifold (wgsl.js:625-977)
  V3<T> gsup(const Geo<T>& g, V3<T> d) const {
    if (g.t == GT_BOX) {
      T th = T(1e-10) * mx(std::fabs(d.x), mx(std::fabs(d.y), std::fabs(d.z)));
      return {d.x >= -th ? g.s.x : -g.s.x, d.y >= -th ? g.s.y : -g.s.y, d.z >= -th ? g.s.z : -g.s.z};
    }
    if (g.t == GT_SPHERE || g.t == GT_CAPSULE) {
      T l2 = dot(d, d);
      V3<T> p = (l2 > T(1e-12) ? d * inverse_sqrt(l2) : V3<T>{1, 0, 0}) * g.s.x;
      if (g.t == GT_CAPSULE) p.z += d.z >= T(0) ? g.s.y : -g.s.y;
      return p;
    }
    if (g.t == GT_CYL) {  // rim point toward d's xy, on the cap d.z points to
      V3<T> xy{d.x, d.y, T(0)};
      T l2 = dot(xy, xy);
      V3<T> p = l2 > T(1e-12) ? xy * inverse_sqrt(l2) * g.s.x : V3<T>{g.s.x, 0, 0};
      if (d.z > T(0)) p.z = g.s.y; else if (d.z < T(0)) p.z = -g.s.y;
      return p;
    }
    T best = T(-1e10);
    V3<T> bp{0, 0, 0};
    for (uint32_t v = 0; v < hnv(g.h); v++) { V3<T> x = hv(g.h, v); T s = dot(x, d); if (s > best) { best = s; bp = x; } }
    return bp;
  }
  V3<T> fsup(const Geo<T>& g, V3<T> d, V3<T> n) const {
    V3<T> p = gsup(g, d);
    if ((g.t == GT_CAPSULE || (g.t == GT_CYL && g.s.z == T(0))) && std::fabs(n.z) <= T(1e-4)) {
      V3<T> rad = normalize(V3<T>{n.x, n.y, 0}) * g.s.x;
      p = {rad.x, rad.y, d.z >= T(0) ? g.s.y : -g.s.y};
    }
    return p;
  }
  static bool planar_geo(const Geo<T>& g) { return g.t == GT_BOX || g.t == GT_CONVEX; }
  static T MEPS() { return T(1e-5); }
  static Trk<T> trk_add(Trk<T> t, V3<T> x, uint32_t id, bool chk) {
    Trk<T> o = t;
    if (id == 0) { o.r = x; o.a2 = 0; o.n = {0, 0, 0}; o.np = false; }
    else if (id == 1) o.p = x;
    else {
      if (chk && o.a2 > T(0)) { T off = dot(o.n, x - o.r); if (off * off > MEPS() * MEPS() * o.a2) o.np = true; }
      V3<T> c = cross(o.p - o.r, x - o.r);
      T a2 = dot(c, c);
      if (a2 > o.a2) { o.a2 = a2; o.n = c; }
      o.p = x;
    }
    return o;
  }
  static bool trk_anchor(const Trk<T>& t, V3<T> x) {
    T off = dot(t.n, x - t.r);
    return !t.np && off * off <= MEPS() * MEPS() * t.a2;
  }
  static V3<T> line_proj_n(V3<T> dir, V3<T> rn) {
    V3<T> nn = cross(cross(dir, rn), dir);
    T l = length(nn);
    return l > T(1e-12) ? nn * (T(1) / l) : rn;
  }
  static void make_proj(const Trk<T>& ta, V3<T> pa, const Trk<T>& tb, V3<T> pb, V3<T> n, Proj<T> out[2]) {
    Proj<T> A{0, {0, 0, 0}}, B{0, {0, 0, 0}};
    if (ta.a2 == T(0) && tb.a2 == T(0)) {
      A.n = line_proj_n(ta.p - ta.r, n); A.d = -dot(T(0.5) * (ta.r + ta.p), A.n);
      B.n = line_proj_n(tb.p - tb.r, n); B.d = -dot(T(0.5) * (tb.r + tb.p), B.n);
      out[0] = A; out[1] = B; return;
    }
    if (ta.a2 > T(0)) { A.n = ta.n * (T(1) / std::sqrt(mx(T(1e-12), ta.a2))); A.d = -dot(pa, A.n); }
    if (tb.a2 > T(0)) { B.n = tb.n * (T(1) / std::sqrt(mx(T(1e-12), tb.a2))); B.d = -dot(pb, B.n); }
    if (ta.a2 == T(0)) { A.n = line_proj_n(ta.p - ta.r, B.n); A.d = -dot(T(0.5) * (ta.r + ta.p), A.n); }
    if (tb.a2 == T(0)) { B.n = line_proj_n(tb.p - tb.r, A.n); B.d = -dot(T(0.5) * (tb.r + tb.p), B.n); }
    out[0] = A; out[1] = B;
  }
  static V3<T> proj_on(const Proj<T>& pr, V3<T> x, V3<T> n) {
    T den = dot(n, pr.n);
    if (std::fabs(den) < T(1e-12)) return x;
    return x + n * (-(dot(x, pr.n) + pr.d) / den);
  }
  static T sarea(V2<T> a, V2<T> b, V2<T> q) { return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x); }
  static bool add_nodup(V2<T>* arr, int cnt, V2<T> v) {
    if (cnt > 0) { V2<T> d = arr[0] - v; if (dot(d, d) < MEPS() * MEPS()) return false; }
    if (cnt > 1) { V2<T> d = arr[cnt - 1] - v; if (dot(d, d) < MEPS() * MEPS()) return false; }
    arr[cnt] = v;
    return true;
  }
  static int trim_in_place(V2<T> s0, V2<T> s1, V2<T>* lp, int cnt0) {
    int cnt = cnt0;
    if (cnt < 3) return cnt;
    V2<T> ia{0, 0}, ib{0, 0};
    int ca = -1, cb = -1;
    bool keep = false;
    bool prev_out = sarea(s0, s1, lp[0]) <= T(0);
    for (int i = 0; i < cnt; i++) {
      int nx = (i + 1) % cnt;
      bool out = sarea(s0, s1, lp[nx]) <= T(0);
      if (out != prev_out) {
        V2<T> a = lp[i], b = lp[nx];
        T sa = sarea(s0, s1, a), sb = sarea(s0, s1, b);
        T t = std::fabs(sa) / std::fabs(sa - sb);
        V2<T> x = (T(1) - t) * a + t * b;
        if (ca < 0) { ca = i; keep = !prev_out; ia = x; } else { cb = i; ib = x; }
      }
      prev_out = out;
    }
    if (ca >= 0 && cb >= 0) {
      int li = -1, ncnt = cnt, i = 0;
      for (int guard = 0; guard < 24 && i < cnt; guard++) {
        if (keep) { li++; lp[li] = lp[i]; }
        if (i == ca || i == cb) {
          V2<T> pt = (i == ca) ? ia : ib;
          if (li == i && !keep) {
            li++;
            for (int k = std::min(ncnt, 9); k > li; k--) lp[k] = lp[k - 1];
            lp[li] = pt;
            ncnt++; i++; cb++; cnt++;
          } else { li++; lp[li] = pt; }
          keep = !keep;
        }
        i++;
      }
      return li + 1;
    }
    return prev_out ? 0 : cnt;
  }
  static int trim_all(V2<T>* tp, int tn, V2<T>* lp, int ln) {
    if (tn <= 1) return std::min(1, ln);
    if (tn == 2 || ln == 2) {
      V2<T> st = lp[0], en = lp[1];
      if (tn == 2) { st = tp[0]; en = tp[1]; }
      V2<T> dir = en - st;
      T l2 = dot(dir, dir);
      if (l2 <= MEPS() * MEPS()) return 0;
      T lo = 0, hi = 1;
      if (tn == 2 && ln == 2) {
        V2<T> os = lp[0], od = lp[1] - os;
        T cr = dir.x * od.y - dir.y * od.x;
        if (std::fabs(cr) > T(1e-6) * std::sqrt(l2 * dot(od, od))) return 0;
        T t0 = dot(os - st, dir) / l2, t1 = dot(lp[1] - st, dir) / l2;
        lo = mx(lo, mn(t0, t1)); hi = mn(hi, mx(t0, t1));
      } else {
        bool useLoop = tn == 2;
        int pc = useLoop ? ln : tn;
        for (int i = 0; i < pc; i++) {
          V2<T> e0 = useLoop ? lp[i] : tp[i];
          V2<T> e1 = useLoop ? lp[(i + 1) % pc] : tp[(i + 1) % pc];
          T a = sarea(e0, e1, st), b = sarea(e0, e1, en);
          if (a < T(0) && b < T(0)) return 0;
          if (a < T(0)) lo = mx(lo, a / (a - b)); else if (b < T(0)) hi = mn(hi, a / (a - b));
        }
      }
      if (hi <= lo) return 0;
      lp[0] = st + lo * dir; lp[1] = st + hi * dir;
      return 2;
    }
    int c = ln;
    V2<T> t0 = tp[0];
    for (int i = 0; i < tn; i++) {
      V2<T> s1 = (i == tn - 1) ? t0 : tp[i + 1];
      c = trim_in_place(tp[i], s1, lp, c);
    }
    return c;
  }
  static int remove_zero_edges(V2<T>* lp, int cnt) {
    if (cnt < 2) return 0;
    int w = 0;
    for (int r = 1; r < cnt; r++) { V2<T> d = lp[r] - lp[w]; if (dot(d, d) > MEPS() * MEPS()) { w++; lp[w] = lp[r]; } }
    int n = w + 1;
    if (w > 0) { V2<T> d = lp[w] - lp[0]; if (dot(d, d) < MEPS() * MEPS()) n = w; }
    return n < 2 ? 0 : n;
  }
  static void calipers(V2<T>* h, int n, int out[4]) {
    int p1 = 0, p3 = 1;
    V2<T> d0 = h[0] - h[1];
    T md = dot(d0, d0);
    const T te = T(1e-3);
    int j = 1;
    for (int i = 0; i < n; i++) {
      V2<T> hi = h[i], hi1 = h[(i + 1) % n];
      for (int g = 0; g < 16; g++) {
        if (sarea(hi, hi1, h[(j + 1) % n]) > sarea(hi, hi1, h[j])) j = (j + 1) % n; else break;
      }
      V2<T> hj = h[j];
      V2<T> a = hi - hj;
      if (dot(a, a) > md * (T(1) + te)) { md = dot(a, a); p1 = i; p3 = j; }
      V2<T> b = hi1 - hj;
      if (dot(b, b) > md * (T(1) + te)) { md = dot(b, b); p1 = (i + 1) % n; p3 = j; }
    }
    int p2 = 0, p4 = 0;
    T m1 = 0, m2 = 0;
    V2<T> q1 = h[p1], q3 = h[p3];
    for (int i = 0; i < n; i++) {
      T ar = sarea(q1, q3, h[i]);
      if (ar > m1 * (T(1) + te)) { m1 = ar; p2 = i; } else if (-ar > m2 * (T(1) + te)) { m2 = -ar; p4 = i; }
    }
    out[0] = p1; out[1] = p2; out[2] = p3; out[3] = p4;
  }
  static void onb(V3<T> n, V3<T> out[2]) {
    if (n.z < T(0)) {
      T a = T(1) / (T(1) - n.z), b = n.x * n.y * a;
      out[0] = {T(1) - n.x * n.x * a, -b, n.x}; out[1] = {b, n.y * n.y * a - T(1), -n.y};
      return;
    }
    T a = T(1) / (T(1) + n.z), b = -n.x * n.y * a;
    out[0] = {T(1) - n.x * n.x * a, b, -n.x}; out[1] = {b, T(1) - n.y * n.y * a, -n.y};
  }
  bool plane_supporting(const Geo<T>& g, V3<T> dir, T d) const {
    T off = dot(dir, gsup(g, dir)) + d;
    return off * off <= MEPS() * MEPS() * dot(dir, dir);
  }
  Man<T> build_manifold(const Geo<T>& ga, const Geo<T>& gb, V4<T> rq, V3<T> rt, V3<T> wa, V3<T> wb, V3<T> n) const {
    const T COS_TILT = T(0.9993908270190958), SIN_TILT = T(0.03489949670250097), COS_DEEP = T(0.9999984769132877);
    static const double PENT[5][2] = {{1.0, 0.0}, {0.30901699437494745, 0.9510565162951535}, {-0.8090169943749473, 0.5877852522924732},
                                      {-0.8090169943749476, -0.587785252292473}, {0.30901699437494723, -0.9510565162951536}};
    Man<T> out;
    out.n = 0;
    V3<T> fna = n;  // cylinder pairs are rejected at packing
    V3<T> tb[2];
    onb(n, tb);
    bool pa = planar_geo(ga), pb = planar_geo(gb);
    V3<T> center = T(0.5) * (wa + wb);
    V2<T> A[10], B[10];
    int ac = 0, bc = 0;
    Trk<T> ta{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, 0, false}, tbk = ta;
    V3<T> lnb = qinvrot(rq, -n), lfb = qinvrot(rq, -fna);
    V3<T> lta = qinvrot(rq, -tb[0]), ltb = qinvrot(rq, -tb[1]);
    for (uint32_t e = 0; e < 5; e++) {
      V2<T> cs = V2<T>{(T)(float)PENT[e][0], (T)(float)PENT[e][1]} * SIN_TILT;
      if (sizeof(T) == 8) cs = V2<T>{(T)PENT[e][0], (T)PENT[e][1]} * SIN_TILT;
      V3<T> x = fsup(ga, n * COS_TILT + cs.x * tb[0] + cs.y * tb[1], fna);
      V3<T> xa = x - center;
      if (add_nodup(A, ac, V2<T>{dot(tb[0], xa), dot(tb[1], xa)})) { ac++; ta = trk_add(ta, x, (uint32_t)(ac - 1), pa); }
      V3<T> y = qrot(rq, fsup(gb, lnb * COS_TILT + cs.x * lta + cs.y * ltb, lfb)) + rt;
      V3<T> yb = y - center;
      if (add_nodup(B, bc, V2<T>{dot(tb[0], yb), dot(tb[1], yb)})) { bc++; tbk = trk_add(tbk, y, (uint32_t)(bc - 1), pb); }
    }
    uint32_t cnt = 0;
    T ndot = 0;
    if (ac >= 2 && bc >= 2 && !(pa && !trk_anchor(ta, wa)) && !(pb && !trk_anchor(tbk, wb))) {
      Proj<T> pr[2];
      make_proj(ta, wa, tbk, wb, n, pr);
      bool ok = std::fabs(dot(n, pr[0].n)) >= COS_TILT && std::fabs(dot(n, pr[1].n)) >= COS_TILT;
      if (ok && pa && (ta.a2 == T(0) || dot(cross(n, pr[0].n), cross(n, pr[0].n)) > T(0))) {
        bool flip = dot(pr[0].n, n) < T(0);
        ok = plane_supporting(ga, flip ? -pr[0].n : pr[0].n, flip ? -pr[0].d : pr[0].d);
      }
      if (ok && pb && (tbk.a2 == T(0) || dot(cross(n, pr[1].n), cross(n, pr[1].n)) > T(0))) {
        bool flip = dot(pr[1].n, n) > T(0);
        V3<T> db = flip ? -pr[1].n : pr[1].n;
        T dd = flip ? -pr[1].d : pr[1].d;
        ok = plane_supporting(gb, qinvrot(rq, db), dd + dot(db, rt));
      }
      if (ok) {
        ndot = std::fabs(dot(pr[0].n, pr[1].n));
        int lc = remove_zero_edges(B, trim_all(A, ac, B, bc));
        if (lc > 1) {
          int idx[4] = {0, 1, 2, 3};
          if (lc > 4) { calipers(B, lc, idx); lc = 4; }
          for (int i = 0; i < lc; i++) {
            V2<T> q = B[idx[i]];
            V3<T> pl = q.x * tb[0] + q.y * tb[1] + center;
            V3<T> a = proj_on(pr[0], pl, n), b = proj_on(pr[1], pl, n);
            out.p[cnt] = T(0.5) * (a + b); out.d[cnt] = dot(b - a, n); cnt++;
          }
        } else ndot = 0;
      }
    }
    if (ndot < COS_DEEP || cnt == 0) { out.p[cnt] = T(0.5) * (wa + wb); out.d[cnt] = dot(wb - wa, n); cnt++; }
    out.n = cnt;
    return out;
  }
  void gen_run(uint32_t sa, uint32_t sb, const Geo<T>& ga, const Geo<T>& gb, V3<T> pa, V4<T> qa, V4<T> rq, V3<T> rt, V3<T> wa, V3<T> wb,
               V3<T> n, T thresh) {
    Man<T> m = build_manifold(ga, gb, rq, rt, wa, wb, n);
    V3<T> nw = qrot(qa, n);
    T ra = (ga.t == GT_SPHERE || ga.t == GT_CAPSULE) ? statF[SF_SHAPE + sa * 7 + 6].y : T(0);
    T rb = (gb.t == GT_SPHERE || gb.t == GT_CAPSULE) ? statF[SF_SHAPE + sb * 7 + 6].y : T(0);
    // A cylinder against a discrete shape rolls when its axis is within 2 degrees of the tangent plane: its
    // points are moved onto the line through the cylinder's center across the axis (wgsl.js:955-967).
    auto discrete = [](uint32_t t) { return t == GT_BOX || t == GT_CONVEX || t == GT_PLANE; };
    bool axA = ga.t == GT_CYL || ga.t == GT_CONE, axB = gb.t == GT_CYL || gb.t == GT_CONE;
    bool roll = false;
    V3<T> rpos{0, 0, 0}, rn{0, 0, 0};
    if ((discrete(ga.t) && axB) || (discrete(gb.t) && axA)) {
      bool onB = discrete(ga.t) && axB;
      V3<T> ax = onB ? qrot(qmul(qa, rq), V3<T>{0, 0, 1}) : qrot(qa, V3<T>{0, 0, 1});
      V3<T> an = onB ? nw : -nw;
      if (std::fabs(dot(ax, an)) <= T(0.03489949670250097)) {
        roll = true;
        rpos = onB ? pa + qrot(qa, rt) : pa;
        rn = normalize(cross(ax, an));
      }
    }
    uint32_t base = hash_u32(sa * 7919u + sb * 104729u);
    for (uint32_t i = 0; i < m.n; i++) {
      V3<T> p = pa + qrot(qa, m.p[i]) + nw * (T(0.5) * ra) - nw * (T(0.5) * rb);
      T d = m.d[i] - ra - rb;
      if (roll) p = p - rn * dot(p - rpos, rn);
      if (d > thresh) continue;
      emit(sa, sb, p, nw, d, hash_u32(base ^ (i * 0x9e3779b1u + 0x51u)));
    }
  }
  T hull_out(uint32_t h, V3<T> x) const {
    T m = T(-1e30);
    for (uint32_t f = 0; f < hnf(h); f++) { V4<T> pl = fplane(hfo(h) + f); m = mx(m, dot(xyz(pl), x) - pl.w); }
    return m;
  }
  V3<T> ray_face(uint32_t h, V3<T> o, V3<T> n, V3<T> s) const {
    T best = T(-2);
    V4<T> pl{0, 0, 0, 0};
    for (uint32_t f = 0; f < hnf(h); f++) { V4<T> q = fplane(hfo(h) + f); if (dot(xyz(q), n) > best) { best = dot(xyz(q), n); pl = q; } }
    V3<T> x = o + n * ((pl.w - dot(xyz(pl), o)) / dot(xyz(pl), n));
    if (std::fabs(dot(x - s, n)) > T(1e-3)) return o + n * dot(s - o, n);
    return x;
  }
  void sup_witness(const Geo<T>& ga, const Geo<T>& gb, V4<T> rq, V3<T> rt, V3<T> n, bool refA, V3<T> w[2]) const {
    V3<T> sA = gsup(ga, n);
    V3<T> lb = gsup(gb, qinvrot(rq, -n));
    V3<T> sB = qrot(rq, lb) + rt;
    V3<T> wb[2] = {ray_face(ga.h, sB, n, sA), sB};
    V3<T> wa[2] = {sA, qrot(rq, ray_face(gb.h, qinvrot(rq, sA - rt), qinvrot(rq, -n), lb)) + rt};
    T ob = mx(hull_out(ga.h, wb[0]), hull_out(gb.h, qinvrot(rq, wb[1] - rt)));
    T oa = mx(hull_out(ga.h, wa[0]), hull_out(gb.h, qinvrot(rq, wa[1] - rt)));
    if ((refA && !(oa < ob - T(1e-6))) || (!refA && ob < oa - T(1e-6))) { w[0] = wb[0]; w[1] = wb[1]; return; }
    w[0] = wa[0]; w[1] = wa[1];
  }
  // ---- MPR for hulls of more than 16 vertices (wgsl.js:310-389, 1206-1214)
  // Support point of hull h along d in the reference's subgroup order: lane l scans vertices l, l + 32, ...
  // keeping its first maximum, and the lowest lane holding the overall maximum wins; serially, lanes in turn
  // with one running strict maximum gives the same point.
  V3<T> sup_lane(uint32_t h, V3<T> d) const {
    T best = T(-1e30);
    V3<T> bp{0, 0, 0};
    for (uint32_t l = 0; l < NLANE; l++)
      for (uint32_t v = l; v < hnv(h); v += NLANE) {
        V3<T> x = hv(h, v);
        T s = dot(d, x);
        if (s > best) { best = s; bp = x; }
      }
    return bp;
  }
  // Support of (B + sphere r) - (A + sphere r) along d, in A's frame.
  V3<T> sup_md(uint32_t ha, uint32_t hb, V4<T> rq, V3<T> t, T r, V3<T> d) const {
    V3<T> sb = qrot(rq, sup_lane(hb, qinvrot(rq, d))) + t;
    V3<T> sa = sup_lane(ha, -d);
    return sb - sa + (T(2) * r) * normalize(d);
  }
  // Minkowski portal refinement (XenoCollide) on the shapes inflated by r: (outward Minkowski normal, inflated
  // depth), w < 0 when the inflated shapes miss.
  V4<T> mpr(uint32_t ha, uint32_t hb, V4<T> rq, V3<T> t, T r) const {
    const V4<T> none{0, 0, 1, -1};
    V3<T> v0 = qrot(rq, xyz(statF[SF_HAABB + hb * 2])) + t - xyz(statF[SF_HAABB + ha * 2]);
    if (dot(v0, v0) < T(1e-14)) v0 = {T(1e-5), 0, 0};
    V3<T> n = -v0;
    V3<T> v1 = sup_md(ha, hb, rq, t, r, n);
    if (dot(v1, n) <= T(0)) return none;
    n = cross(v1, v0);
    if (dot(n, n) < T(1e-14)) {
      V3<T> nn = normalize(v1 - v0);
      return v4(nn, dot(v1, nn));
    }
    V3<T> v2 = sup_md(ha, hb, rq, t, r, n);
    if (dot(v2, n) <= T(0)) return none;
    n = cross(v1 - v0, v2 - v0);
    if (dot(n, v0) > T(0)) { V3<T> tmp = v1; v1 = v2; v2 = tmp; n = -n; }
    V3<T> v3{0, 0, 0};
    bool found = false;
    for (uint32_t i = 0; i < 32; i++) {
      v3 = sup_md(ha, hb, rq, t, r, n);
      if (dot(v3, n) <= T(0)) return none;
      if (dot(cross(v1, v3), v0) < T(0)) { v2 = v3; n = cross(v1 - v0, v3 - v0); continue; }
      if (dot(cross(v3, v2), v0) < T(0)) { v1 = v3; n = cross(v3 - v0, v2 - v0); continue; }
      found = true;
      break;
    }
    if (!found) return none;
    for (uint32_t i = 0; i < 32; i++) {
      n = cross(v2 - v1, v3 - v1);
      T ln = length(n);
      if (ln < T(1e-12)) break;
      n = n / ln;
      if (dot(n, v1 - v0) < T(0)) n = -n;
      V3<T> v4_ = sup_md(ha, hb, rq, t, r, n);
      T d1 = dot(n, v1);
      if (dot(v4_, n) - d1 < T(1e-6) || i == 31) {
        if (d1 < T(0)) return none;
        return v4(n, d1);
      }
      if (dot(v4_, n) < T(0)) return none;
      V3<T> c = cross(v4_, v0);
      if (dot(v1, c) > T(0)) { if (dot(v2, c) > T(0)) v1 = v4_; else v3 = v4_; }
      else { if (dot(v3, c) > T(0)) v2 = v4_; else v1 = v4_; }
    }
    T d1 = dot(n, v1);
    if (d1 < T(0)) return none;
    return v4(n, d1);
  }
  // Large hulls: MPR's normal, the reference side from face alignment.
  void mpr_generic(uint32_t sa, uint32_t sb, uint32_t ha, uint32_t hb, V3<T> pa, V4<T> qa, V4<T> rBA, V3<T> tBA, T thresh, V3<T> n) {
    T alA = T(-2);
    for (uint32_t f = 0; f < hnf(ha); f++) alA = mx(alA, dot(xyz(fplane(hfo(ha) + f)), n));
    T alB = T(-2);
    for (uint32_t f = 0; f < hnf(hb); f++) alB = mx(alB, -dot(qrot(rBA, xyz(fplane(hfo(hb) + f))), n));
    Geo<T> ga = geo(sa), gb = geo(sb);
    V3<T> w[2];
    sup_witness(ga, gb, rBA, tBA, n, alA >= alB, w);
    gen_run(sa, sb, ga, gb, pa, qa, rBA, tBA, w[0], w[1], n, thresh);
  }
  void sat_generic(uint32_t sa, uint32_t sb, uint32_t ha, uint32_t hb, V3<T> pa, V4<T> qa, V4<T> rBA, V3<T> tBA, T thresh, FaceQ<T> fa,
                   FaceQ<T> fb, const EdgeQ<T>& eq) {
    const T SAT_TIE = T(1e-5);
    Geo<T> ga = geo(sa), gb = geo(sb);
    if (eq.sep > mx(fa.sep, fb.sep) + SAT_TIE) {
      V3<T> c[2];
      closest_seg(eq.a0, eq.a1, eq.b0, eq.b1, c);
      gen_run(sa, sb, ga, gb, pa, qa, rBA, tBA, c[0], c[1], eq.axis, thresh);
      return;
    }
    bool refA = !(fb.sep > fa.sep + SAT_TIE);
    V3<T> n = refA ? xyz(fplane(hfo(ha) + fa.face)) : -qrot(rBA, xyz(fplane(hfo(hb) + fb.face)));
    V3<T> w[2];
    sup_witness(ga, gb, rBA, tBA, n, refA, w);
    gen_run(sa, sb, ga, gb, pa, qa, rBA, tBA, w[0], w[1], n, thresh);
  }
  void plane_convex(uint32_t sa, uint32_t sb, V3<T> pa, V4<T> qa, V3<T> pb, V4<T> qb, T thresh) {
    Geo<T> gb = geo(sb);
    V3<T> nw = qrot(qa, V3<T>{0, 0, 1});
    T L = T(10) * (length(xyz(statF[SF_HAABB + gb.h * 2])) + length(xyz(statF[SF_HAABB + gb.h * 2 + 1])));
    V3<T> c = pb - nw * (dot(pb - pa, nw) + L);
    V4<T> rq = qmul(qconj(qa), qb);
    V3<T> rt = qinvrot(qa, pb - c);
    V3<T> n{0, 0, 1};
    V3<T> wb = qrot(rq, gsup(gb, qinvrot(rq, -n))) + rt;
    T sep = wb.z - L;
    if (sep > thresh) return;
    Geo<T> box{GT_BOX, {L, L, L}, 0};
    gen_run(sa, sb, box, gb, c, qa, rq, rt, wb - n * sep, wb, n, thresh);
  }
  void capsule_generic(uint32_t c, uint32_t h, V3<T> pc, V4<T> qc, V3<T> wp, V4<T> wq, V3<T> wc, V3<T> wh, V3<T> nh, T thresh) {
    if (dot(wh - wc, nh) - statF[SF_SHAPE + c * 7 + 6].y > thresh) return;
    V4<T> rq = qmul(qconj(qc), wq);
    V3<T> rt = qinvrot(qc, wp - pc);
    gen_run(c, h, geo(c), geo(h), pc, qc, rq, rt, qrot(rq, wc) + rt, qrot(rq, wh) + rt, qrot(rq, nh), thresh);
  }
  void narrow_pair(uint32_t pi) {  // wgsl.js:1072-1195
    uint32_t x0 = pairs[pi * 2], y0 = pairs[pi * 2 + 1];
    bool swp = geo_type(x0) > geo_type(y0);
    uint32_t sa = swp ? y0 : x0, sb = swp ? x0 : y0;
    uint32_t ha = tu(statF[SF_SHAPE + sa * 7].y), hb = tu(statF[SF_SHAPE + sb * 7].y);
    V3<T> pa = xyz(poses[P_SHAPE + sa * 2]);
    V4<T> qa = poses[P_SHAPE + sa * 2 + 1];
    V3<T> pb = xyz(poses[P_SHAPE + sb * 2]);
    V4<T> qb = poses[P_SHAPE + sb * 2 + 1];
    T thresh = mn(statF[SF_SHAPE + sa * 7].w + statF[SF_SHAPE + sb * 7].w, F_GAP_GATE);
    V4<T> rBA = qmul(qconj(qa), qb);
    V3<T> tBA = qinvrot(qa, pb - pa);
    V4<T> rAB = qconj(rBA);
    V3<T> tAB = qinvrot(qb, pa - pb);
    V4<T> ka = shape_prim(sa), kb = shape_prim(sb);
    if (ka.x == T(4) || kb.x == T(4)) {
      if (geo_type(sb) == GT_CONVEX) plane_convex(sa, sb, pa, qa, pb, qb, thresh);
      else plane_contacts(sa, sb, pa, qa, pb, qb, thresh, false);
      return;
    }
    if (ka.x != T(0) && kb.x != T(0)) { prim_prim(sa, sb, pa, qa, pb, qb, thresh); return; }
    if (ka.x != T(0) || kb.x != T(0)) {
      bool flip = ka.x != T(0);
      uint32_t hs = flip ? sb : sa, ps = flip ? sa : sb, h = flip ? hb : ha;
      V4<T> pr = flip ? ka : kb;
      V4<T> rq = flip ? rAB : rBA;
      V3<T> rt = flip ? tAB : tBA;
      V3<T> u = qrot(rq, V3<T>{0, 0, pr.z});
      V3<T> e0 = rt - u, e1 = rt + u;
      FaceQ<T> fq = reduce_lanes_max([&](uint32_t l) { return face_query_seg(h, e0, e1, pr.y, l); });
      if (fq.sep > thresh) return;
      V3<T> wp = flip ? pb : pa;
      V4<T> wq = flip ? qb : qa;
      if (fq.sep + pr.y > T(1e-6) || (pr.x == T(2) && !seg_hits_hull(h, e0, e1))) {
        SegQ<T> r[NLANE];
        for (uint32_t l = 0; l < NLANE; l++) r[l] = seg_hull_query(h, e0, e1, l);
        uint32_t b = 0;
        for (uint32_t k = 1; k < NLANE; k++) if (r[k].d2 < r[b].d2) b = k;
        SegQ<T> q = r[b];
        if (pr.x == T(2)) {
          T d = std::sqrt(q.d2);
          if (d >= T(1e-9)) {
            V3<T> nh = (q.ph - q.ps) / d;
            capsule_generic(ps, hs, flip ? pa : pb, flip ? qa : qb, wp, wq, q.ps + nh * T(1e-4), q.ph, nh, thresh);
          }
        } else {
          T d = std::sqrt(q.d2);
          T phi = d - pr.y;
          V3<T> n = (q.ps - q.ph) / mx(d, T(1e-9));
          if (phi > thresh || d < T(1e-9)) {
          } else if (dot(n, xyz(fplane(hfo(h) + fq.face))) > T(0.999)) {
            seg_on_face(hs, ps, h, fq.face, e0, e1, pr.y, wp, wq, thresh, flip);
          } else {
            uint32_t sa2 = flip ? ps : hs, sb2 = flip ? hs : ps;
            V3<T> nw = qrot(wq, n);
            uint32_t key = hash_u32(hash_u32(sa2 * 7919u + sb2 * 104729u) ^ (q.face * 131u + q.kind * 17u + 0x9e37u));
            emit(sa2, sb2, wp + qrot(wq, q.ph + n * (T(0.5) * phi)), flip ? -nw : nw, phi, key);
          }
        }
        return;
      }
      if (pr.x == T(2)) {
        V4<T> pl = fplane(hfo(h) + fq.face);
        V3<T> x = dot(xyz(pl), e0) <= dot(xyz(pl), e1) ? e0 : e1;
        capsule_generic(ps, hs, flip ? pa : pb, flip ? qa : qb, wp, wq, x - xyz(pl) * T(1e-4), x - xyz(pl) * (dot(xyz(pl), x) - pl.w), -xyz(pl),
                        thresh);
      } else {
        seg_on_face(hs, ps, h, fq.face, e0, e1, pr.y, wp, wq, thresh, flip);
      }
      return;
    }
    FaceQ<T> fa = reduce_lanes_max([&](uint32_t l) { return face_query(ha, hb, rBA, tBA, l); });
    if (fa.sep > thresh) return;
    FaceQ<T> fb = reduce_lanes_max([&](uint32_t l) { return face_query(hb, ha, rAB, tAB, l); });
    if (fb.sep > thresh) return;
    if (!(hnv(ha) <= 16 && hnv(hb) <= 16)) {
      // Face-only SAT misjudges edge contacts on large hulls: MPR's normal instead, unless a face axis
      // separates the hulls better (wgsl.js:1164-1176).
      V4<T> m = mpr(ha, hb, rBA, tBA, T(0.5) * thresh);
      if (m.w >= T(0)) {
        V3<T> n = -xyz(m);
        T sep = dot(n, tBA) + support_min(hb, qinvrot(rBA, n)) - support_max(ha, n);
        if (mx(fa.sep, fb.sep) > sep + T(1e-4)) {
          EdgeQ<T> none{T(-1e30), {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, 0};
          sat_generic(sa, sb, ha, hb, pa, qa, rBA, tBA, thresh, fa, fb, none);
        } else {
          mpr_generic(sa, sb, ha, hb, pa, qa, rBA, tBA, thresh, n);
        }
      }
      return;
    }
    EdgeQ<T> r[NLANE];
    for (uint32_t l = 0; l < NLANE; l++) r[l] = edge_query(ha, hb, rBA, tBA, l);
    uint32_t b = 0;
    for (uint32_t k = 1; k < NLANE; k++) if (r[k].sep > r[b].sep) b = k;
    EdgeQ<T> eq = r[b];
    if (eq.sep > thresh) return;
    sat_generic(sa, sb, ha, hb, pa, qa, rBA, tBA, thresh, fa, fb, eq);
  }
  void narrow() {
    craw.clear();
    counters[1] = 0;
    uint32_t np = std::min(counters[0], MAXP);
    for (uint32_t p = 0; p < np; p++) narrow_pair(p);
  }

  // ============================ contact sort (wgsl.js:3895-3980) ============================
  void csort() {
    uint32_t nc = std::min(counters[1], MAXC);
    // (b + 1, a + 1) in place of the 16-bit packed pair (a + 1) | (b + 1) << 16: the same order, for any body count
    struct K { uint32_t k[5]; };
    std::vector<K> keys(nc);
    for (uint32_t c = 0; c < nc; c++) {
      const Contact<T>& r = craw[c];
      uint32_t pa = (uint32_t)(ti(r.q[2].x) + 1), pb = (uint32_t)(ti(r.q[2].y) + 1);
      uint32_t feat = tu(r.q[2].z) | (tu(r.q[2].w) << 16);
      uint32_t bits = hash_u32(hash_u32(hash_u32(fbits(r.q[0].x)) ^ fbits(r.q[0].y)) ^ fbits(r.q[0].z)) ^ fbits(r.q[0].w);
      keys[c] = {{pb, pa, feat, bits, c}};
    }
    std::sort(keys.begin(), keys.end(), [](const K& a, const K& b) {
      for (int i = 0; i < 5; i++) if (a.k[i] != b.k[i]) return a.k[i] < b.k[i];
      return false;
    });
    contacts.resize(nc);
    for (uint32_t i = 0; i < nc; i++) contacts[i] = craw[keys[i].k[4]];
  }

  // ============================ articulation (artwgsl.js) ============================
  void art_fk(uint32_t a) {  // artwgsl.js:98-187
    uint32_t ai[4];
    art_info(a, ai);
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k;
      uint32_t o = SF_LINK + l * 6;
      V3<T> pp{0, 0, 0};
      V4<T> pq{0, 0, 0, 1};
      int par = link_parent(l);
      if (par >= 0) { pp = ld3(lo((uint32_t)par) + L_POS); pq = link_rot((uint32_t)par); }
      V3<T> jp = pp + qrot(pq, xyz(statF[o + 1]));
      V4<T> jq = qmul(pq, statF[o + 2]);
      V3<T> axis = xyz(statF[o + 5]);
      V4<T> mq{0, 0, 0, 1};
      V3<T> mp{0, 0, 0};
      uint32_t t = link_type(l), qs = link_qs(l);
      if (t == JT_REVOLUTE) mq = qaxis(axis, art[A_Q + qs]);
      if (t == JT_PRISMATIC) mp = axis * art[A_Q + qs];
      if (t == JT_BALL) mq = normalize4(V4<T>{art[A_Q + qs], art[A_Q + qs + 1], art[A_Q + qs + 2], art[A_Q + qs + 3]});
      // D6: translations along the leading axes, then rotations about axes carried through the rotations before
      // them (artwgsl.js:117-141)
      uint32_t d0 = link_dof(l), nlin = link_nlin(l), nang = link_ndof(l) - nlin;
      V3<T> ang[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      if (t == JT_D6) {
        for (uint32_t c = 0; c < nlin; c++) mp += dof_axis(d0 + c) * art[A_Q + qs + c];
        uint32_t iq = qs + nlin;
        if (nang == 1) {
          ang[0] = dof_axis(d0 + nlin);
          mq = qaxis(ang[0], art[A_Q + iq]);
        }
        if (nang == 2) {
          V3<T> x0 = normalize(dof_axis(d0 + nlin));
          V3<T> x2 = normalize(cross(x0, dof_axis(d0 + nlin + 1)));
          ang[0] = x0;
          V4<T> q0 = qaxis(x0, art[A_Q + iq]);
          ang[1] = qrot(q0, cross(x2, x0));
          mq = qmul(qaxis(ang[1], art[A_Q + iq + 1]), q0);
        }
        if (nang == 3) {
          ang[0] = dof_axis(d0 + nlin);
          V4<T> q0 = qaxis(ang[0], art[A_Q + iq]);
          ang[1] = qrot(q0, dof_axis(d0 + nlin + 1));
          V4<T> q1 = qaxis(ang[1], art[A_Q + iq + 1]);
          ang[2] = qrot(qmul(q1, q0), dof_axis(d0 + nlin + 2));
          mq = qmul(qaxis(ang[2], art[A_Q + iq + 2]), qmul(q1, q0));
        }
      }
      if (t == JT_FREE) {
        mp = {art[A_Q + qs], art[A_Q + qs + 1], art[A_Q + qs + 2]};
        mq = normalize4(V4<T>{art[A_Q + qs + 3], art[A_Q + qs + 4], art[A_Q + qs + 5], art[A_Q + qs + 6]});
      }
      V4<T> cq = qmul(jq, mq);
      V3<T> cp = jp + qrot(jq, mp);
      V4<T> iq = qconj(statF[o + 4]);
      V4<T> lq = normalize4(qmul(cq, iq));
      V3<T> lp = cp + qrot(cq, -qrot(iq, xyz(statF[o + 3])));
      uint32_t L = lo(l);
      st3(L + L_POS, lp); st3(L + L_ROT, xyz(lq)); art[L + L_ROT + 3] = lq.w;
      V3<T> com = lp + qrot(lq, body_com_local(link_body(l)));
      st3(L + L_COM, com);
      V3<T> aw = qrot(jq, axis);
      V3<T> rj = com - jp;
      for (uint32_t c = 0; c < JM; c++) stsv(L + L_S + c * 6, sv0<T>());
      if (t == JT_REVOLUTE) stsv(L + L_S, SV<T>{aw, cross(aw, rj)});
      if (t == JT_PRISMATIC) stsv(L + L_S, SV<T>{{0, 0, 0}, aw});
      if (t == JT_BALL)
        for (uint32_t c = 0; c < 3; c++) {
          V3<T> e{0, 0, 0}; comp(e, c) = 1;
          V3<T> ew = qrot(jq, e);
          stsv(L + L_S + c * 6, SV<T>{ew, cross(ew, rj)});
        }
      V3<T> rf = com - cp;
      st3(L + L_JR, rf);
      if (t == JT_D6) {
        for (uint32_t c = 0; c < nlin; c++) stsv(L + L_S + c * 6, SV<T>{{0, 0, 0}, qrot(jq, dof_axis(d0 + c))});
        for (uint32_t c = 0; c < nang; c++) {
          V3<T> ew = qrot(jq, ang[c]);
          stsv(L + L_S + (nlin + c) * 6, SV<T>{ew, cross(ew, rf)});
        }
      }
      if (t == JT_FREE)
        for (uint32_t c = 0; c < 3; c++) {
          V3<T> e{0, 0, 0}; comp(e, c) = 1;
          V3<T> ew = qrot(jq, e);
          stsv(L + L_S + c * 6, SV<T>{{0, 0, 0}, ew});
          stsv(L + L_S + (c + 3) * 6, SV<T>{ew, cross(ew, rf)});
        }
    }
  }
  void inv_spd(T* D, uint32_t n, T* Di) const {  // artwgsl.js:192-207
    for (uint32_t i = 0; i < JM * JM; i++) Di[i] = 0;
    for (uint32_t i = 0; i < n; i++) Di[i * JM + i] = 1;
    for (uint32_t c = 0; c < n; c++) {
      T pv = T(1) / mx(D[c * JM + c], T(1e-12));
      for (uint32_t k = 0; k < n; k++) { D[c * JM + k] *= pv; Di[c * JM + k] *= pv; }
      for (uint32_t r = 0; r < n; r++) {
        if (r == c) continue;
        T f = D[r * JM + c];
        if (f == T(0)) continue;
        for (uint32_t k = 0; k < n; k++) { D[r * JM + k] -= f * D[c * JM + k]; Di[r * JM + k] -= f * Di[c * JM + k]; }
      }
    }
  }
  void add_shifted_inertia(uint32_t o, const M3<T>& A, const M3<T>& B, const M3<T>& C, V3<T> r) {
    M3<T> R = skew3(r);
    stm(o, msub(madd(msub(madd(ldm(o), A), mul(B, R)), mul(R, transpose(B))), mul(mul(R, C), R)));
    stm(o + 9, madd(madd(ldm(o + 9), B), mul(R, C)));
    stm(o + 18, madd(ldm(o + 18), C));
  }
  void art_factor(uint32_t a) {  // artwgsl.js:210-369
    uint32_t ai[4];
    art_info(a, ai);
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k, L = lo(l);
      int par = link_parent(l);
      SV<T> vp = sv0<T>();
      if (par >= 0) vp = mshift(ldsv(lo((uint32_t)par) + L_V), link_r(l));
      SV<T> sq = sv0<T>();
      for (uint32_t c = 0; c < link_ndof(l); c++) sq = sadd(sq, sscale(ldS(l, c), art[A_QD + link_dof(l) + c]));
      SV<T> v = sadd(vp, sq);
      stsv(L + L_V, v);
      SV<T> cb = crm(vp, sq);
      uint32_t jt = link_type(l);
      if (jt == JT_D6) {
        // rotation axes transported by earlier D6 angles: d(a_k)/dt = sum_{j<k} a_j x a_k qd_j (artwgsl.js:223-233)
        uint32_t nlin = link_nlin(l);
        V3<T> ca{0, 0, 0};
        for (uint32_t j = nlin; j < link_ndof(l); j++)
          for (uint32_t k2 = j + 1; k2 < link_ndof(l); k2++)
            ca += cross(ldS(l, j).a, ldS(l, k2).a) * art[A_QD + link_dof(l) + j] * art[A_QD + link_dof(l) + k2];
        cb = sadd(cb, SV<T>{ca, cross(ca, ld3(L + L_JR))});
      }
      if (jt == JT_FREE || jt == JT_D6) {
        uint32_t nlin = jt == JT_FREE ? 3 : link_nlin(l);
        V3<T> vo{0, 0, 0}, wj{0, 0, 0};
        for (uint32_t c = 0; c < link_ndof(l); c++) {
          SV<T> sc = ldS(l, c);
          T q = art[A_QD + link_dof(l) + c];
          if (c < nlin) vo += sc.l * q; else wj += sc.a * q;
        }
        cb = sadd(cb, SV<T>{{0, 0, 0}, -cross(wj, vo)});
      }
      stsv(L + L_C, cb);
      int b = link_body(l);
      T m = body_mass(b);
      M3<T> R = qmat(link_rot(l));
      M3<T> Ic = mul(mul(R, body_inertia(b)), transpose(R));
      stm(L + L_IA, Ic);
      stm(L + L_IA + 9, m3zero<T>());
      M3<T> Mm; Mm.c[0] = {m, 0, 0}; Mm.c[1] = {0, m, 0}; Mm.c[2] = {0, 0, m};
      stm(L + L_IA + 18, Mm);
      SV<T> pa0 = crf(v, imul(L + L_IA, v));
      if ((tu(statF[SF_BODY + (uint32_t)b * 8 + 7].w) & 2u) != 0) pa0 = sadd(pa0, SV<T>{{0, 0, 0}, m * g});
      pa0 = sadd(pa0, SV<T>{-ld3(A_BF + (uint32_t)b * 8 + 4), -ld3(A_BF + (uint32_t)b * 8)});
      stsv(L + L_PA, pa0);
    }
    for (uint32_t kk = 0; kk < ai[1]; kk++) {
      uint32_t l = ai[0] + ai[1] - 1 - kk, L = lo(l);
      uint32_t n = link_ndof(l);
      SV<T> pA = ldsv(L + L_PA);
      T D[36], u[6] = {0, 0, 0, 0, 0, 0};
      for (uint32_t i = 0; i < JM * JM; i++) D[i] = 0;
      for (uint32_t c = 0; c < n; c++) stsv(L + L_U + c * 6, imul(L + L_IA, ldS(l, c)));
      for (uint32_t r = 0; r < n; r++) {
        uint32_t d = link_dof(l) + r;
        for (uint32_t c = 0; c < n; c++) D[r * JM + c] = sdot(ldS(l, r), ldU(l, c));
        V4<T> dr = dof_drive(d);
        D[r * JM + r] += dr.x;
        T damp = dof_passive(d).x;
        D[r * JM + r] += dt * damp;
        T tau = -damp * art[A_QD + d] + art[A_JF + d];
        // Passive spring, explicit (artwgsl.js:298-300).
        V4<T> sp = statF[SF_DOF + d * 5 + 4];
        if (sp.x != T(0)) tau += sp.x * (sp.y - art[A_Q + link_qs(l) + r]);
        // Implicit PD drive toward (A_TGT, A_TGTV), saturated on the implicit torque (artwgsl.js:301-313).
        if (!DRIVE_ROWS && (dr.y > T(0) || dr.z > T(0))) {
          T q = art[A_Q + link_qs(l) + r], qd = art[A_QD + d];
          T err = art[A_TGT + d] - q;
          T tv = art[A_TGTV + d];
          T m = mx(D[r * JM + r], T(1e-9));
          T qdn = (m * qd + dt * (dr.y * err + dr.z * tv)) / (m + dt * dr.z + dt * dt * dr.y);
          T ti = dr.y * (err - dt * qdn) + dr.z * (tv - qdn);
          if (std::fabs(ti) <= dr.w || dr.w <= T(0)) {
            D[r * JM + r] += dt * dt * dr.y + dt * dr.z;
            tau += dr.y * err - dt * dr.y * qd + dr.z * (tv - qd);
          } else {
            tau += std::min(std::max(ti, -dr.w), dr.w);
          }
        }
        u[r] = tau - sdot(ldS(l, r), pA);
      }
      T Di[36];
      inv_spd(D, n, Di);
      for (uint32_t i = 0; i < JM * JM; i++) art[L + L_DI + i] = Di[i];
      for (uint32_t r = 0; r < JM; r++) art[L + L_UU + r] = u[r];
      int par = link_parent(l);
      if (par < 0) continue;
      M3<T> A = ldm(L + L_IA), B = ldm(L + L_IA + 9), C = ldm(L + L_IA + 18);
      SV<T> Duu = sv0<T>();
      for (uint32_t r = 0; r < n; r++) {
        SV<T> Ur = ldU(l, r);
        T du = 0;
        for (uint32_t c = 0; c < n; c++) {
          SV<T> Uc = ldU(l, c);
          T w = ldDI(l, r, c);
          A = msub(A, mscale(w, outer3(Ur.a, Uc.a)));
          B = msub(B, mscale(w, outer3(Ur.a, Uc.l)));
          C = msub(C, mscale(w, outer3(Ur.l, Uc.l)));
          du += w * u[c];
        }
        Duu = sadd(Duu, sscale(Ur, du));
      }
      stm(L + L_IA, A); stm(L + L_IA + 9, B); stm(L + L_IA + 18, C);
      SV<T> cbias = ldsv(L + L_C);
      SV<T> pa = sadd(sadd(pA, imul(L + L_IA, cbias)), Duu);
      uint32_t P0 = lo((uint32_t)par);
      V3<T> r = link_r(l);
      add_shifted_inertia(P0 + L_IA, A, B, C, r);
      stsv(P0 + L_PA, sadd(ldsv(P0 + L_PA), fshift(pa, r)));
    }
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k, L = lo(l);
      int par = link_parent(l);
      SV<T> ap{{0, 0, 0}, -g};
      if (par >= 0) ap = mshift(ldsv(lo((uint32_t)par) + L_A), link_r(l));
      SV<T> a1 = sadd(ap, ldsv(L + L_C));
      uint32_t n = link_ndof(l);
      SV<T> acc = a1;
      for (uint32_t r = 0; r < n; r++) {
        T qdd = 0;
        for (uint32_t c = 0; c < n; c++) qdd += ldDI(l, r, c) * (art[L + L_UU + c] - sdot(ldU(l, c), a1));
        uint32_t d = link_dof(l) + r;
        art[A_QDH + d] = art[A_QD + d] + dt * qdd;
        acc = sadd(acc, sscale(ldS(l, r), qdd));
      }
      stsv(L + L_A, acc);
      if (JFL) {
        // Load-dependent joint friction: the joint's transmitted force and torque under the free dynamics
        // (articulated force at the child) raise its friction bound (artwgsl.js:361-367).
        SV<T> f = sadd(imul(L + L_IA, acc), ldsv(L + L_PA));
        T fm = length(f.l) + length(f.a);
        for (uint32_t r = 0; r < n; r++) { uint32_t d = link_dof(l) + r; art[A_JFL + d] = statF[SF_DOF + d * 5 + 3].w * fm; }
      }
    }
  }
  uint32_t link_path(uint32_t l0, uint32_t* path) const {
    uint32_t n = 0;
    int l = (int)l0;
    while (!(l < 0 || n >= MAXPATH)) { path[n++] = (uint32_t)l; l = link_parent((uint32_t)l); }
    return n;
  }
  T path_test_dqd(uint32_t tl, uint32_t td, T tau, const uint32_t* path, uint32_t np) const {  // artwgsl.js:375-414 with F = 0
    std::vector<T> ub(MAXPATH * JM, T(0));
    SV<T> p = sscale(sv0<T>(), T(-1));
    for (uint32_t i = 0; i < np; i++) {
      uint32_t l = path[i], n = link_ndof(l);
      T u[6] = {0, 0, 0, 0, 0, 0};
      for (uint32_t r = 0; r < n; r++) {
        T t = 0;
        if (l == tl && r == td) t = tau;
        u[r] = t - sdot(ldS(l, r), p);
        ub[i * JM + r] = u[r];
      }
      for (uint32_t r = 0; r < n; r++) {
        T du = 0;
        for (uint32_t c = 0; c < n; c++) du += ldDI(l, r, c) * u[c];
        p = sadd(p, sscale(ldU(l, r), du));
      }
      p = fshift(p, link_r(l));
    }
    SV<T> dv = sv0<T>();
    T out = 0;
    for (uint32_t ii = 0; ii < np; ii++) {
      uint32_t i = np - 1 - ii, l = path[i], n = link_ndof(l);
      dv = mshift(dv, link_r(l));
      T dq[6];
      for (uint32_t r = 0; r < n; r++) {
        dq[r] = 0;
        for (uint32_t c = 0; c < n; c++) dq[r] += ldDI(l, r, c) * (ub[i * JM + c] - sdot(ldU(l, c), dv));
      }
      for (uint32_t r = 0; r < n; r++) {
        dv = sadd(dv, sscale(ldS(l, r), dq[r]));
        if (l == tl && r == td) out = dq[r];
      }
    }
    return out;
  }
  T dof_hinv(uint32_t l, uint32_t r) const {
    std::vector<uint32_t> path(MAXPATH);
    uint32_t np = link_path(l, path.data());
    return path_test_dqd(l, r, T(1), path.data(), np);
  }
  void link_response(uint32_t l, const uint32_t* split_count) {  // artwgsl.js:431-510
    std::vector<uint32_t> path(MAXPATH);
    uint32_t np = link_path(l, path.data());
    uint32_t PD = MAXPATH * JM;
    std::vector<T> ub(6 * PD, T(0));
    SV<T> p[6];
    for (uint32_t e = 0; e < 6; e++) {
      V3<T> f{0, 0, 0}, t{0, 0, 0};
      if (e < 3) comp(f, e) = 1; else comp(t, e - 3) = 1;
      p[e] = {-t, -f};
    }
    for (uint32_t i = 0; i < np; i++) {
      uint32_t pl = path[i], n = link_ndof(pl);
      SV<T> S[6], U[6];
      T DI[36];
      for (uint32_t c = 0; c < n; c++) { S[c] = ldS(pl, c); U[c] = ldU(pl, c); }
      for (uint32_t x = 0; x < n * JM; x++) DI[x] = art[lo(pl) + L_DI + x];
      V3<T> r = link_r(pl);
      for (uint32_t e = 0; e < 6; e++) {
        T u[6];
        for (uint32_t c = 0; c < n; c++) { u[c] = T(0) - sdot(S[c], p[e]); ub[e * PD + i * JM + c] = u[c]; }
        SV<T> pe = p[e];
        for (uint32_t rr = 0; rr < n; rr++) {
          T du = 0;
          for (uint32_t c = 0; c < n; c++) du += DI[rr * JM + c] * u[c];
          pe = sadd(pe, sscale(U[rr], du));
        }
        p[e] = fshift(pe, r);
      }
    }
    SV<T> dv[6];
    for (uint32_t e = 0; e < 6; e++) dv[e] = sv0<T>();
    for (uint32_t ii = 0; ii < np; ii++) {
      uint32_t i = np - 1 - ii, pl = path[i], n = link_ndof(pl);
      SV<T> S[6], U[6];
      T DI[36];
      for (uint32_t c = 0; c < n; c++) { S[c] = ldS(pl, c); U[c] = ldU(pl, c); }
      for (uint32_t x = 0; x < n * JM; x++) DI[x] = art[lo(pl) + L_DI + x];
      V3<T> r = link_r(pl);
      for (uint32_t e = 0; e < 6; e++) {
        SV<T> d = mshift(dv[e], r);
        T dq[6];
        for (uint32_t rr = 0; rr < n; rr++) {
          dq[rr] = 0;
          for (uint32_t c = 0; c < n; c++) dq[rr] += DI[rr * JM + c] * (ub[e * PD + i * JM + c] - sdot(U[c], d));
        }
        for (uint32_t rr = 0; rr < n; rr++) d = sadd(d, sscale(S[rr], dq[rr]));
        dv[e] = d;
      }
    }
    T W[36];
    for (uint32_t e = 0; e < 6; e++) {
      W[0 * 6 + e] = dv[e].l.x; W[1 * 6 + e] = dv[e].l.y; W[2 * 6 + e] = dv[e].l.z;
      W[3 * 6 + e] = dv[e].a.x; W[4 * 6 + e] = dv[e].a.y; W[5 * 6 + e] = dv[e].a.z;
    }
    for (uint32_t k = 0; k < 36; k++) art[lo(l) + L_W2 + k] = W[k];
    for (uint32_t i = 0; i < np; i++) {
      T extra = (T)split_count[i] - T(1);
      if (extra <= T(0)) continue;
      uint32_t pl = path[i], n = link_ndof(pl);
      T DI[36];
      for (uint32_t x = 0; x < n * JM; x++) DI[x] = art[lo(pl) + L_DI + x];
      T t[36];
      for (uint32_t cc = 0; cc < 6; cc++)
        for (uint32_t x = 0; x < n; x++) {
          T v = 0;
          for (uint32_t y = 0; y < n; y++) v += DI[x * JM + y] * ub[cc * PD + i * JM + y];
          t[cc * 6 + x] = v;
        }
      for (uint32_t rr = 0; rr < 6; rr++)
        for (uint32_t cc = 0; cc < 6; cc++) {
          T sv = 0;
          for (uint32_t x = 0; x < n; x++) sv += ub[rr * PD + i * JM + x] * t[cc * 6 + x];
          W[rr * 6 + cc] += extra * sv;
        }
    }
    for (uint32_t k = 0; k < 36; k++) art[lo(l) + L_W + k] = W[k];
  }
  // Dense rows of the loop closures: three per closure, five for a revolute one (wgsl.js:2367-2368).
  uint32_t loop_rows() const {
    uint32_t n = 0;
    for (uint32_t k = 0; k < NLOOP; k++) n += statF[SF_LOOP + k * 5].w > T(0.5) ? 2 : 1;
    return n;
  }
  // Angular Jacobian rows d0, d1 of link x's world angular velocity over its root path (artwgsl.js:542-554).
  void add_gen_ang(uint32_t x, T sgn, V3<T> d0, V3<T> d1, uint32_t ds, T* J) const {
    int l = (int)x;
    while (l >= 0) {
      uint32_t k = (uint32_t)l;
      for (uint32_t r = 0; r < link_ndof(k); r++) {
        V3<T> wv = ldS(k, r).a;
        uint32_t i = link_dof(k) + r - ds;
        J[i] += sgn * dot(d0, wv); J[NDA + i] += sgn * dot(d1, wv);
      }
      l = link_parent(k);
    }
  }
  void add_gen_jac(uint32_t x, V3<T> p, T sgn, V3<T> d0, V3<T> d1, V3<T> d2, uint32_t ds, T* J) const {
    int l = (int)x;
    while (l >= 0) {
      uint32_t k = (uint32_t)l;
      V3<T> c = ld3(lo(k) + L_COM);
      for (uint32_t r = 0; r < link_ndof(k); r++) {
        SV<T> S = ldS(k, r);
        V3<T> v = S.l + cross(S.a, p - c);
        uint32_t i = link_dof(k) + r - ds;
        J[i] += sgn * dot(d0, v); J[NDA + i] += sgn * dot(d1, v); J[2 * NDA + i] += sgn * dot(d2, v);
      }
      l = link_parent(k);
    }
  }
  void art_solve_gen(uint32_t a, const T* tau, T* out) const {  // artwgsl.js:557-596
    uint32_t ai[4];
    art_info(a, ai);
    std::vector<SV<T>> p(MAXLP), dvs(MAXLP);
    std::vector<T> uu(MAXLP * 6, T(0));
    for (uint32_t k = 0; k < ai[1]; k++) p[k] = sv0<T>();
    for (uint32_t kk = 0; kk < ai[1]; kk++) {
      uint32_t k = ai[1] - 1 - kk, l = ai[0] + k, n = link_ndof(l);
      T u[6] = {0, 0, 0, 0, 0, 0};
      for (uint32_t r = 0; r < n; r++) u[r] = tau[link_dof(l) + r - ai[2]] - sdot(ldS(l, r), p[k]);
      for (uint32_t r = 0; r < 6; r++) uu[k * 6 + r] = u[r];
      int par = link_parent(l);
      if (par >= 0) {
        SV<T> pp = p[k];
        for (uint32_t r = 0; r < n; r++) {
          T du = 0;
          for (uint32_t c = 0; c < n; c++) du += ldDI(l, r, c) * u[c];
          pp = sadd(pp, sscale(ldU(l, r), du));
        }
        uint32_t pk = (uint32_t)par - ai[0];
        p[pk] = sadd(p[pk], fshift(pp, link_r(l)));
      }
    }
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k;
      int par = link_parent(l);
      SV<T> dv = sv0<T>();
      if (par >= 0) dv = mshift(dvs[(uint32_t)par - ai[0]], link_r(l));
      uint32_t n = link_ndof(l);
      T dq[6];
      for (uint32_t r = 0; r < n; r++) {
        dq[r] = 0;
        for (uint32_t c = 0; c < n; c++) dq[r] += ldDI(l, r, c) * (uu[k * 6 + c] - sdot(ldU(l, c), dv));
      }
      for (uint32_t r = 0; r < n; r++) { dv = sadd(dv, sscale(ldS(l, r), dq[r])); out[link_dof(l) + r - ai[2]] = dq[r]; }
      dvs[k] = dv;
    }
  }

  // ============================ prepare (wgsl.js:1787-2494) ============================
  void body_response(int b, V3<T> f, V3<T> t, V3<T> out[2]) const {  // wgsl.js:1312-1328
    if (b < 0) { out[0] = {0, 0, 0}; out[1] = {0, 0, 0}; return; }
    if (body_kind(b) == 1) {
      V4<T> q = poses[P_BODY + (uint32_t)b * 2 + 1];
      M3<T> R = qmat(q);
      out[0] = f * body_inv_mass(b);
      out[1] = mul(R, mul(body_inv_inertia(b), mul(transpose(R), t)));
      return;
    }
    uint32_t wo = lo((uint32_t)body_link(b)) + L_W;
    T x[6] = {f.x, f.y, f.z, t.x, t.y, t.z}, o[6];
    for (uint32_t r = 0; r < 6; r++) { T s = 0; for (uint32_t c = 0; c < 6; c++) s += art[wo + r * 6 + c] * x[c]; o[r] = s; }
    out[0] = {o[0], o[1], o[2]}; out[1] = {o[3], o[4], o[5]};
  }
  void body_response_unsplit(int b, V3<T> f, V3<T> t, V3<T> out[2]) const {
    if (b < 0 || body_kind(b) != 2) { body_response(b, f, t, out); return; }
    uint32_t wo = lo((uint32_t)body_link(b)) + L_W2;
    T x[6] = {f.x, f.y, f.z, t.x, t.y, t.z}, o[6];
    for (uint32_t r = 0; r < 6; r++) { T s = 0; for (uint32_t c = 0; c < 6; c++) s += art[wo + r * 6 + c] * x[c]; o[r] = s; }
    out[0] = {o[0], o[1], o[2]}; out[1] = {o[3], o[4], o[5]};
  }
  bool restitution_fires(T e, T phi, T vn) const {
    if (REST == 0 || e <= T(0) || vn >= -F_REST_SPEED) return false;
    return phi <= T(1e-6) || phi + dt * vn <= T(1e-6);
  }
  void pre_twist(int b, V3<T> out[2]) const {
    out[0] = {0, 0, 0}; out[1] = {0, 0, 0};
    if (b < 0) return;
    if (body_kind(b) == 1) { out[0] = xyz(vel[(uint32_t)b * 2]); out[1] = xyz(vel[(uint32_t)b * 2 + 1]); return; }
    if (body_kind(b) == 2) { uint32_t o = lo((uint32_t)body_link(b)) + L_V; out[0] = ld3(o + 3); out[1] = ld3(o); }
  }
  T row_weight(T phi) const { if (phi > T(0) || reg <= T(0)) return T(1); return T(1) / (T(1) + reg); }
  static uint64_t pair_key(int a, int b) { return ((uint64_t)(uint32_t)(b + 1) << 32) | (uint32_t)(a + 1); }
  static int key_a(uint64_t k) { return (int)(uint32_t)k - 1; }
  static int key_b(uint64_t k) { return (int)(uint32_t)(k >> 32) - 1; }
  uint64_t contact_pair(uint32_t c) const {
    const V4<T>& m = contacts[c].q[2];
    return pair_key(ti(m.x), ti(m.y));
  }
  uint32_t unit_of(uint32_t p, uint32_t nu) const {
    uint32_t lo_ = 0, hi = nu;
    for (uint32_t it = 0; it < 32 && hi - lo_ > 1; it++) { uint32_t mid = (lo_ + hi) / 2; if (S_USTART[mid] <= p) lo_ = mid; else hi = mid; }
    return lo_;
  }
  struct BPC { uint32_t k[3]; V2<T> pos; T gap; uint32_t pk; };
  BPC bp_contact(uint32_t c) const {  // wgsl.js:1475-1495
    static const float BINS[20][3] = {
        {-0.57095543f, 0.75026579f, 0.33333337f}, {-0.64404845f, 0.17218052f, 0.74535593f}, {-0.86918069f, -0.36525859f, 0.33333330f},
        {-0.93522697f, -0.11932883f, -0.33333339f}, {-0.75091368f, 0.57010308f, -0.33333336f}, {0.36427171f, 0.86959472f, -0.33333337f},
        {-0.17291149f, 0.64385240f, -0.74535609f}, {0.47113684f, 0.47167205f, 0.74535599f}, {0.11826704f, 0.93536185f, 0.33333335f},
        {0.17291155f, -0.64385245f, 0.74535603f}, {0.0f, 0.0f, 1.0f}, {-0.11826707f, -0.93536184f, -0.33333335f},
        {-0.36427169f, -0.86959473f, 0.33333337f}, {0.0f, 0.0f, -1.0f}, {-0.47113677f, -0.47167196f, -0.74535609f},
        {0.93522695f, 0.11932898f, 0.33333339f}, {0.75091367f, -0.57010306f, 0.33333340f}, {0.57095545f, -0.75026579f, -0.33333334f},
        {0.64404840f, -0.17218051f, -0.74535597f}, {0.86918063f, 0.36525872f, -0.33333332f}};
    const Contact<T>& C = contacts[c];
    V4<T> c0 = C.q[0], c1 = C.q[1], cm = C.q[2], c4 = C.q[4];
    int a = ti(cm.x), b = ti(cm.y);
    uint32_t ga = tu(c4.z), gb = tu(c4.w);
    V3<T> ng = gb < ga ? -xyz(c1) : xyz(c1);
    int rb = a >= 0 ? a : b;
    V3<T> org{0, 0, 0};
    if (rb >= 0) { V4<T> q = poses[P_BODY + (uint32_t)rb * 2 + 1]; org = xyz(poses[P_BODY + (uint32_t)rb * 2]) - qrot(q, body_com_local(rb)); }
    uint32_t bin = 0;
    T bd = T(-2);
    for (uint32_t i = 0; i < 20; i++) {
      V3<T> B{(T)BINS[i][0], (T)BINS[i][1], (T)BINS[i][2]};
      T d = dot(ng, B);
      if (d > bd) { bd = d; bin = i; }
    }
    V3<T> fn{(T)BINS[bin][0], (T)BINS[bin][1], (T)BINS[bin][2]};
    V3<T> rv = std::fabs(fn.y) < T(0.9) ? V3<T>{0, 1, 0} : V3<T>{1, 0, 0};
    V3<T> u = normalize(rv - dot(rv, fn) * fn);
    V3<T> v = cross(fn, u);
    V3<T> x = xyz(c0) - org;
    V2<T> pos{dot(x, u), dot(x, v)};
    uint32_t cx = (uint32_t)((int32_t)clampv(std::floor(pos.x / F_BP_CELL + T(0.5)), T(-31), T(31)) + 31);
    uint32_t cy = (uint32_t)((int32_t)clampv(std::floor(pos.y / F_BP_CELL + T(0.5)), T(-31), T(31)) + 31);
    uint32_t qx = (uint32_t)((int32_t)clampv(std::floor(pos.x * T(2000)), T(-32768), T(32767)) + 32768);
    uint32_t qy = (uint32_t)((int32_t)clampv(std::floor(pos.y * T(2000)), T(-8192), T(8191)) + 8192);
    BPC r;
    r.k[0] = std::min(ga, gb); r.k[1] = std::max(ga, gb); r.k[2] = bin | (cx << 5) | (cy << 11);
    r.pos = pos; r.gap = c0.w; r.pk = (qx << 15) | (qy << 1) | 1u;
    return r;
  }
  void bp_value(const BPC& g, uint32_t q, uint32_t out[2]) const {  // wgsl.js:1498-1503
    if (q == 6) { out[0] = fflip_u(fbits(-g.gap)); out[1] = g.pk; return; }
    if (q > 6 && g.gap > T(0)) { out[0] = 0; out[1] = 0; return; }
    uint32_t k = q > 6 ? q - 7 : q;
    T cs, sn;
    if (sizeof(T) == 4) { float ang = (float)k * 1.0471975511965976f; cs = (T)cosf_c(ang); sn = (T)sinf_c(ang); }
    else { double ang = (double)k * 1.0471975511965976; cs = (T)std::cos(ang); sn = (T)std::sin(ang); }
    out[0] = fflip_u(fbits(dot(g.pos, V2<T>{cs, sn})));
    out[1] = g.pk;
  }
  uint32_t first_free_color(int a, int b) {  // wgsl.js:1233-1246
    for (uint32_t w = 0; w < NCOLW; w++) {
      uint32_t used = 0;
      if (a >= 0) used |= wmask[(uint32_t)a * NCOLW + w];
      if (b >= 0) used |= wmask[(uint32_t)b * NCOLW + w];
      if (used != 0xffffffffu) {
        uint32_t k = ctz32(~used);
        if (a >= 0) wmask[(uint32_t)a * NCOLW + w] |= 1u << k;
        if (b >= 0) wmask[(uint32_t)b * NCOLW + w] |= 1u << k;
        return w * 32 + k;
      }
    }
    return NCOL;
  }
  uint32_t anchor_hash(uint32_t pair, uint32_t k) const { return std::max(hash_u32(pair * 2u + k), 1u); }
  // Contact torsion (wgsl.js:1526-1557): one spin row per articulated unit about the normal of its deepest point,
  // over the members touching in that plane (normal within 0.995, offset along it within 0.002, penetrating),
  // with the largest member torsion radius r.  Header in the unit's first row: (Ma lin, mu), (Ma ang, 1 / K),
  // (Mb lin, spin impulse), (Mb ang, r), (n, first row's member flag); every row's TO + 4 .w is its member flag.
  void build_torsion(uint32_t sp) {
    uint32_t st = S_USTART[sp];
    uint32_t n = std::min(S_USIZE[sp], MAXC - std::min(st, MAXC));
    if (n == 0) return;
    uint64_t pair = S_UKEY[sp];
    int a = key_a(pair), b = key_b(pair);
    uint32_t L = st * ROWS;
    bool art_u = (a >= 0 && body_kind(a) >= 2) || (b >= 0 && body_kind(b) >= 2);
    uint32_t seed = st;
    for (uint32_t i = st; i < st + n; i++) if (rows[i * ROWS + 1].w < rows[seed * ROWS + 1].w) seed = i;
    V3<T> ns = xyz(rows[seed * ROWS]), ps = xyz(rows[seed * ROWS + TO + 4]);
    T r = 0;
    for (uint32_t i = st; i < st + n; i++) {
      uint32_t o = i * ROWS;
      bool member = art_u && rows[o + 6].w <= T(0) && dot(xyz(rows[o]), ns) >= T(0.995) &&
                    std::fabs(dot(xyz(rows[o + TO + 4]) - ps, ns)) <= T(0.002);
      if (member) r = mx(r, rows[o + TO + 3].w);
      rows[o + TO + 4].w = member ? T(1) : T(0);
    }
    T mu = rows[seed * ROWS + 2].w;
    if (r <= T(0) || mu <= T(0)) { rows[L + TO] = {0, 0, 0, 0}; return; }
    V3<T> Ma[2], Mb[2];
    body_response(a, V3<T>{0, 0, 0}, -ns, Ma);
    body_response(b, V3<T>{0, 0, 0}, ns, Mb);
    T k = dot(-ns, Ma[1]) + dot(ns, Mb[1]);
    T lead = rows[L + TO + 4].w;
    rows[L + TO] = v4(Ma[0], mu);
    rows[L + TO + 1] = v4(Ma[1], k > T(1e-12) ? T(1) / (k + F_CFM) : T(0));
    rows[L + TO + 2] = v4(Mb[0], T(0));
    rows[L + TO + 3] = v4(Mb[1], r);
    rows[L + TO + 4] = v4(ns, lead);
  }
  void build_patch(uint32_t sp) {  // wgsl.js:1562-1725
    uint32_t st = S_USTART[sp];
    uint32_t n = std::min(S_USIZE[sp], MAXC - std::min(st, MAXC));
    if (n == 0) return;
    uint64_t key = S_UKEY[sp];
    int a = key_a(key), b = key_b(key);
    // Anchor-table pair id: the reference's (a + 1) | (b + 1) << 16, taken modulo 2^32 past 65535 bodies.
    uint32_t pair = (uint32_t)(a + 1) | ((uint32_t)(b + 1) << 16);
    V3<T> ca{0, 0, 0}, cb{0, 0, 0};
    V4<T> qa{0, 0, 0, 1};
    if (a >= 0) { ca = xyz(poses[P_BODY + (uint32_t)a * 2]); qa = poses[P_BODY + (uint32_t)a * 2 + 1]; }
    if (b >= 0) cb = xyz(poses[P_BODY + (uint32_t)b * 2]);
    T r = mx(mn(body_radius(a), body_radius(b)), T(1e-4));
    uint32_t seed = st;
    for (uint32_t i = st; i < st + n; i++) if (rows[i * ROWS + 1].w < rows[seed * ROWS + 1].w) seed = i;
    V3<T> ns = xyz(rows[seed * ROWS]);
    uint32_t kept = S_OSTART[sp] - st;  // index into S_KEPT: kept + i
    auto cmu = [&](uint32_t i) { const Contact<T>& C = contacts[S_KEPT[kept + i]]; return V2<T>{C.smu[0], C.smu[1]}; };
    V2<T> mu_seed = cmu(seed);
    uint32_t first = seed;
    for (uint32_t i = st; i < st + n; i++) {
      uint32_t o = i * ROWS;
      V3<T> dseed = xyz(rows[o + PO + 2]) - xyz(rows[seed * ROWS + PO + 2]);
      V2<T> mi = cmu(i);
      if (dot(xyz(rows[o]), ns) < T(0.995) || mi.x != mu_seed.x || mi.y != mu_seed.y || std::fabs(dot(dseed, ns)) > T(0.02) * r ||
          length(dseed) > T(2) * r)
        continue;
      rows[o + PO].w = 0;
      V3<T> p = xyz(rows[o + PO + 2]), pf = xyz(rows[first * ROWS + PO + 2]);
      if (p.x < pf.x || (p.x == pf.x && (p.y < pf.y || (p.y == pf.y && p.z < pf.z)))) first = i;
    }
    V3<T> pf = xyz(rows[first * ROWS + PO + 2]);
    uint32_t second = first;
    T far = T(1e-8) * r * r;
    for (uint32_t i = st; i < st + n; i++) {
      if (rows[i * ROWS + PO].w != T(0)) continue;
      V3<T> d = xyz(rows[i * ROWS + PO + 2]) - pf;
      if (dot(d, d) > far) { far = dot(d, d); second = i; }
    }
    uint32_t na = second != first ? 2 : 1;
    V3<T> loc0 = xyz(rows[first * ROWS + PO + 2]), loc1 = xyz(rows[second * ROWS + PO + 2]);
    if (na == 2) {
      V3<T> t0 = cross(ns, V3<T>{1, 0, 0});
      if (dot(t0, t0) < T(1e-6)) t0 = cross(ns, V3<T>{0, 1, 0});
      t0 = normalize(t0);
      V3<T> t1 = normalize(cross(ns, t0));
      T sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0, cnt = 0;
      V3<T> cen{0, 0, 0};
      for (uint32_t i = st; i < st + n; i++) {
        if (rows[i * ROWS + PO].w != T(0)) continue;
        V3<T> pc = xyz(rows[i * ROWS + PO + 2]);
        V3<T> d = pc - pf;
        T x = dot(d, t0), y = dot(d, t1);
        sx += x; sy += y; sxx += x * x; sxy += x * y; syy += y * y; cen += pc; cnt += T(1);
      }
      T xx = sxx - sx * sx / cnt, xy = sxy - sx * sy / cnt, yy = syy - sy * sy / cnt;
      T spread = std::sqrt((xx - yy) * (xx - yy) + T(4) * xy * xy);
      if (spread > T(1e-6) * (xx + yy)) {
        T big = T(0.5) * (xx + yy + spread);
        V2<T> a2{xy, big - xx};
        if (xx >= yy) a2 = {big - yy, xy};
        a2 = normalize(a2);
        V3<T> ax = a2.x * t0 + a2.y * t1;
        T lo_ = T(1e30), hi = T(-1e30);
        for (uint32_t i = st; i < st + n; i++) {
          if (rows[i * ROWS + PO].w != T(0)) continue;
          T pr = dot(xyz(rows[i * ROWS + PO + 2]) - pf, ax);
          if (pr < lo_) { lo_ = pr; first = i; }
          if (pr > hi) { hi = pr; second = i; }
        }
        V3<T> centroid = cen / cnt;
        loc0 = centroid + dot(xyz(rows[first * ROWS + PO + 2]) - centroid, ax) * ax;
        loc1 = centroid + dot(xyz(rows[second * ROWS + PO + 2]) - centroid, ax) * ax;
        V3<T> across = normalize(cross(ns, ax));
        T tol = mx(T(1e-5) * r, T(0.05) * (hi - lo_));
        V3<T> e0{0, 0, 0}, e1{0, 0, 0};
        T c0 = 0, c1 = 0;
        for (uint32_t i = st; i < st + n; i++) {
          if (rows[i * ROWS + PO].w != T(0)) continue;
          V3<T> pc = xyz(rows[i * ROWS + PO + 2]);
          T pr = dot(pc - pf, ax);
          if (pr <= lo_ + tol) { e0 += pc; c0 += T(1); }
          if (pr >= hi - tol) { e1 += pc; c1 += T(1); }
        }
        e0 = e0 / mx(c0, T(1)); e1 = e1 / mx(c1, T(1));
        loc0 += dot(e0 - centroid, across) * across + dot(e0 - centroid, ns) * ns;
        loc1 += dot(e1 - centroid, across) * across + dot(e1 - centroid, ns) * ns;
      }
    }
    T load = 0;
    for (uint32_t i = st; i < st + n; i++) {
      if (rows[i * ROWS + PO].w != T(0)) continue;
      load += rows[i * ROWS + 6].x;
      if (i != first && !(na == 2 && i == second)) { V4<T> l6 = rows[i * ROWS + 6]; rows[i * ROWS + 6] = {l6.x, 0, 0, l6.w}; }
    }
    uint32_t rbase = parity * AH;
    for (uint32_t k = 0; k < na; k++) {
      uint32_t i = k == 1 ? second : first, o = i * ROWS;
      V3<T> loc = k == 1 ? loc1 : loc0;
      V3<T> nn = xyz(rows[o]), t1 = xyz(rows[o + 1]), t2 = xyz(rows[o + 2]);
      V3<T> fra = loc - ca, frb = loc - cb;
      V3<T> Ma[2][2], Mb[2][2];
      V3<T> dirs[2] = {t1, t2};
      for (uint32_t q = 0; q < 2; q++) {
        body_response(a, -dirs[q], -cross(fra, dirs[q]), Ma[q]);
        body_response(b, dirs[q], cross(frb, dirs[q]), Mb[q]);
        rows[o + 11 + q * 4] = v4(Ma[q][0], T(0)); rows[o + 12 + q * 4] = v4(Ma[q][1], T(0));
        rows[o + 13 + q * 4] = v4(Mb[q][0], T(0)); rows[o + 14 + q * 4] = v4(Mb[q][1], T(0));
      }
      T K[4];
      for (uint32_t x = 0; x < 2; x++)
        for (uint32_t y = 0; y < 2; y++) {
          V3<T> d = dirs[x];
          K[x * 2 + y] = dot(-d, Ma[y][0]) + dot(-cross(fra, d), Ma[y][1]) + dot(d, Mb[y][0]) + dot(cross(frb, d), Mb[y][1]);
        }
      T ks = mx(K[0] + K[3], T(1e-30));
      T k00 = K[0] + T(1e-6) * ks, k01 = T(0.5) * (K[1] + K[2]), k11 = K[3] + T(1e-6) * ks;
      T det = k00 * k11 - k01 * k01;
      rows[o + 5] = {k11 / det, -k01 / det, k00 / det, rows[o + 5].w};
      rows[o + 2] = v4(t2, rows[o + 2].w / (T)na);
      V3<T> e{0, 0, 0}, impv{0, 0, 0};
      bool carried = false;
      T best = T(4) * r * r;
      for (uint32_t kk = 0; kk < 2; kk++) {
        uint32_t h = anchor_hash(pair, kk) % AH;
        for (uint32_t probe = 0; probe < 32; probe++) {
          uint32_t key = akey[rbase + h];
          if (key == 0) break;
          const AnchorVal<T>& av = aval[rbase + h];
          if (key == anchor_hash(pair, kk) && av.pair == pair && av.k == kk) {
            V3<T> prev = ca + qrot(qa, av.ll);
            V3<T> dl = prev - loc;
            if (dot(dl, dl) <= best && std::fabs(dot(dl, nn)) <= T(0.05) * r) {
              V3<T> ew = qrot(qa, av.el);
              V3<T> on = qrot(qa, av.nl);
              T cs = dot(on, nn);
              if (cs > T(0)) { V3<T> ax = cross(on, nn); V3<T> ce = cross(ax, ew); ew += ce + cross(ax, ce) / (T(1) + cs); }
              ew -= nn * dot(nn, ew);
              if (dot(ew, ew) <= T(0.01) * r * r) { e = ew; best = dot(dl, dl); carried = true; impv = qrot(qa, av.il); }
            }
            break;
          }
          h = (h + 1) % AH;
        }
      }
      // Warm start: the carried anchor impulse, else the contact's, within the anchor's share mu / na of the
      // members' warm normal load (wgsl.js:1711-1719).
      if (warm_scale > T(0)) {
        V2<T> lw{rows[o + 6].y, rows[o + 6].z};
        if (carried) lw = V2<T>{dot(impv, t1), dot(impv, t2)} * warm_scale;
        T rad = mx(rows[o + 2].w * load, T(0));
        if (length(lw) > rad) lw = lw * (rad / length(lw));
        rows[o + 6] = {rows[o + 6].x, lw.x, lw.y, rows[o + 6].w};
      }
      rows[o + PO] = v4(fra, T(2));
      rows[o + PO + 1] = v4(frb, (T)k);
      rows[o + PO + 2] = v4(e, T(0));
      rows[o + PO + 3] = v4(qinvrot(qa, fra), put_u32((T*)nullptr, pair));
    }
  }
  // ---- sleeping (wgsl.js:1370-1457; artwgsl.js:41-49): nodes are free bodies (b) and articulations (NB + a)
  int body_node(int b) const {
    if (b < 0) return -1;
    uint32_t k = body_kind(b);
    if (k == 1) return b;
    if (k == 2 || k == 3) return (int)NB + (int)body_art(b);
    return -1;
  }
  bool node_asleep(int n) const { return SLEEP && n >= 0 && art[A_SLP + (uint32_t)n * 4 + 1] > T(0.5); }
  bool pair_asleep(int a, int b) const {
    if (!SLEEP) return false;
    int na = body_node(a), nb = body_node(b);
    return (na < 0 || node_asleep(na)) && (nb < 0 || node_asleep(nb));
  }
  // Islands of nodes joined by contacts, each rooted at its lowest node; island flags on the root: 1 supported
  // (a fixed-base articulation, or a contact with static geometry at depth <= 1e-5), 2 vetoed (moving, or an
  // articulation with drives), 4 a member still too young to sleep.  Ages grow while an island is supported and
  // quiet; an island sleeps once all its members reached the quiet time.
  void sleep_islands() {
    const uint32_t NN = NB + NA;
    for (uint32_t n = 0; n < NN; n++) { uf[n] = n; uff[n] = 0; }
    S_META[5] = 0;
    uint32_t nc = std::min(counters[1], MAXC);
    auto find = [&](uint32_t n) { while (uf[n] != n) n = uf[n]; return n; };
    for (uint32_t c = 0; c < nc; c++) {
      const V4<T>& m = contacts[c].q[2];
      int na = body_node(ti(m.x)), nb = body_node(ti(m.y));
      if ((na < 0 || nb < 0) && std::max(na, nb) >= 0 && contacts[c].q[0].w <= T(1e-5)) uff[(uint32_t)std::max(na, nb)] |= 1u;
    }
    for (uint32_t c = 0; c < nc; c++) {
      const V4<T>& m = contacts[c].q[2];
      int na = body_node(ti(m.x)), nb = body_node(ti(m.y));
      if (na < 0 || nb < 0) continue;
      uint32_t ra = find((uint32_t)na), rb = find((uint32_t)nb);
      if (ra != rb) uf[std::max(ra, rb)] = std::min(ra, rb);
    }
    for (uint32_t n = 0; n < NN; n++) uf[n] = find(n);
    for (uint32_t n = 0; n < NN; n++) {
      bool moving = false, fixed = false, veto = false;
      if (n < NB) {
        if (body_kind((int)n) != 1) continue;
        if (art[A_SLP + n * 4 + 1] < T(0.5)) moving = length(xyz(vel[n * 2])) > F_SLEEP_LIN || length(xyz(vel[n * 2 + 1])) > F_SLEEP_ANG;
      } else {
        uint32_t ai[4];
        art_info(n - NB, ai);
        fixed = link_type(ai[0]) != JT_FREE;
        for (uint32_t k = 0; k < ai[1]; k++) {
          SV<T> v = ldsv(lo(ai[0] + k) + L_V);
          if (art[A_SLP + n * 4 + 1] < T(0.5) && (length(v.l) > F_SLEEP_LIN || length(v.a) > F_SLEEP_ANG)) moving = true;
        }
        for (uint32_t d = ai[2]; d < ai[2] + ai[3]; d++) { V4<T> dr = dof_drive(d); if (dr.y > T(0) || dr.z > T(0)) veto = true; }
        if (statU[SU_ACON + n - NB] != 0) veto = true;  // owns a mimic
      }
      uint32_t f = (fixed ? 1u : 0u) | ((moving || veto) ? 2u : 0u) | (uff[n] & 1u);
      uff[uf[n]] |= f;
    }
    for (uint32_t n = 0; n < NN; n++) {
      if (n < NB && body_kind((int)n) != 1) continue;
      uint32_t r = uf[n], f = uff[r];
      T age = art[A_SLP + n * 4];
      if ((f & 2u) != 0 || (f & 1u) == 0) age = T(0); else age = mn(age + dt, F_SLEEP_TIME);
      art[A_SLP + n * 4] = age;
      if (age < F_SLEEP_TIME) uff[r] |= 4u;
    }
    for (uint32_t n = 0; n < NN; n++) {
      if (n < NB && body_kind((int)n) != 1) continue;
      bool s = (uff[uf[n]] & 4u) == 0;
      art[A_SLP + n * 4 + 1] = s ? T(1) : T(0);
      if (s) S_META[5]++;
    }
  }
  // ---- contact warm start (wgsl.js:2175-2190, 2421-2435; 3614-3631, 3737-3767)
  // The previous substep's impulses stored under `key`, scaled by the warm scale; the tangent part is projected on
  // (t1, t2) and kept within the friction cone mu * ln.
  void warm_lookup(uint32_t key, T mu, V3<T> t1, V3<T> t2, T& ln, V2<T>& lt) const {
    if (!WH) return;
    V3<T> lf{0, 0, 0};
    uint32_t rbase = parity * WH, h = key % WH;
    for (uint32_t probe = 0; probe < 32; probe++) {
      const WarmSlot<T>& s = wtab[rbase + h];
      if (s.key == 0) break;
      if (s.key == key) { ln = s.v.x * warm_scale; lf = V3<T>{s.v.y, s.v.z, s.v.w} * warm_scale; break; }
      h = (h + 1) % WH;
    }
    lt = {dot(lf, t1), dot(lf, t2)};
    T lim = mu * ln, lm = length(lt);
    if (lm > lim) lt = lt * (lim / mx(lm, T(1e-12)));
  }
  // Row i's key (contact rows, then dense rows; 0xffffffff for a bilateral row, which stores nothing).
  uint32_t warm_key(uint32_t i, uint32_t nc) const {
    if (i < nc) return (uint32_t)rows[i * ROWS + 3].w | ((uint32_t)rows[i * ROWS + 4].w << 16);
    uint32_t o = A_DENSE + (i - nc) * DPS;
    if (art[o + 21] > T(0.5)) return 0xffffffffu;
    return (uint32_t)art[o + 1] | ((uint32_t)art[o + 2] << 16);
  }
  V4<T> warm_value(uint32_t i, uint32_t nc) const {
    if (i < nc) {
      uint32_t o = i * ROWS;
      V4<T> lam = rows[o + 6];
      V3<T> f = xyz(rows[o + 1]) * lam.y + xyz(rows[o + 2]) * lam.z;
      return {lam.x, f.x, f.y, f.z};
    }
    uint32_t o = A_DENSE + (i - nc) * DPS;
    V3<T> f = ld3(o + 12) * art[o + 10] + ld3(o + 15) * art[o + 11];
    return {art[o + 9], f.x, f.y, f.z};
  }
  // Table of the next substep: every key inserted by linear probing (at most 32 probes); its value is that of
  // the row with the largest normal impulse (negative and -0 count as 0), ties going to the lowest row.
  void warm_store(uint32_t nc) {
    uint32_t nw = nc + std::min(S_META[3], MAXDENSE), wb = (1 - parity) * WH;
    std::vector<int64_t> best(WH, -1);
    for (uint32_t i = 0; i < nw; i++) {
      uint32_t key = warm_key(i, nc);
      if (key == 0xffffffffu) continue;
      uint32_t h = key % WH;
      for (uint32_t probe = 0; probe < 32; probe++) {
        WarmSlot<T>& s = wtab[wb + h];
        if (s.key == 0) s.key = key;
        if (s.key == key) {
          T x = warm_value(i, nc).x, y = best[h] < 0 ? T(0) : warm_value((uint32_t)best[h], nc).x;
          if (best[h] < 0 || (x > T(0) ? x : T(0)) > (y > T(0) ? y : T(0))) best[h] = i;
          break;
        }
        h = (h + 1) % WH;
      }
    }
    for (uint32_t h = 0; h < WH; h++)
      if (best[h] >= 0) wtab[wb + h].v = warm_value((uint32_t)best[h], nc);
  }
  void prepare() {
    uint32_t ROWSN = ROWS;
    // Phase 0
    std::fill(S_HKEY.begin(), S_HKEY.end(), 0ull);
    std::fill(S_HCNT.begin(), S_HCNT.end(), 0u);
    if (PATCH) for (uint32_t i = 0; i < AH; i++) akey[(1 - parity) * AH + i] = 0;
    for (uint32_t i = 0; i < WH; i++) wtab[(1 - parity) * WH + i].key = 0;
    std::fill(wlinkhit.begin(), wlinkhit.end(), 0u);
    std::fill(wmask.begin(), wmask.end(), 0u);
    std::fill(wcount.begin(), wcount.end(), 0u);
    S_META[3] = 0;
    std::fill(wjsplit.begin(), wjsplit.end(), 0u);
    if (SLEEP) sleep_islands();
    // Recorded target paths (60 Hz samples), interpolated at every substep at its midpoint half a frame back
    // (or held per frame), plus the integral term of the first dofs (wgsl.js:1800-1820); off once the user
    // set targets.
    uint32_t driven = (control & 1u) ? 0u : DRIVEN;
    if (driven && PATH_FRAMES) {
      T ph = T((float)phase);  // the per-substep parameter is a float
      for (uint32_t i = 0; i < driven; i++) {
        T ramp = (ph + (T(substep_in_frame) + T(0.5)) * dt) * T(60) - T(0.5);
        T frame = mn(mx(TGT_HELD ? ph * T(60) : ramp, T(0)), T(PATH_FRAMES - 1u));
        uint32_t fa = (uint32_t)frame, fb = std::min(fa + 1u, PATH_FRAMES - 1u);
        uint32_t ia = SF_PATH * 4 + fa * driven + i, ib = SF_PATH * 4 + fb * driven + i;
        T pa = (&statF[ia / 4].x)[ia % 4], pb = (&statF[ib / 4].x)[ib % 4];
        T t = frame - T(fa);
        T value = pa * (T(1) - t) + pb * t;
        if (i < INTEGRAL_COUNT) {
          if (substep_in_frame == 0) {
            T err = mn(mx(value - art[A_Q + i], T(-0.08)), T(0.08));
            art[A_INT + i] = mn(mx(art[A_INT + i] + T(3) * FRAME_DT * err, T(-0.12)), T(0.12));
          }
          value += art[A_INT + i];
        }
        art[A_TGT + i] = value;
      }
    }
    // Demo controller, every substep (update_control, wgsl.js:1821-1825): the fader sets the wheel's target speed.
    if (CTRL_KIND == 1) {
      V4<T> cc = statF[SF_CTRL];
      art[A_TGTV + CTRL_B] = cc.y * mn(mx(art[A_Q + CTRL_A] / cc.x, T(0)), T(1));
    }
    for (uint32_t a = 0; a < NA; a++) {
      uint32_t ai[4];
      art_info(a, ai);
      if (VLIM_ON) {
        // Velocity-limit prescale: the articulation's scalar dof velocities scaled so the most exceeded limit
        // is met; then the dofs near their limit (activation fraction) and whether any is (wgsl.js:1828-1855).
        T ratio = T(1);
        for (uint32_t k = 0; k < ai[3]; k++) {
          uint32_t d = ai[2] + k;
          T vl = dof_passive(d).z;
          if (vl < T(0) || fused_dof(d)) continue;
          T s = std::fabs(art[A_QD + d]);
          if (s > vl) ratio = std::min(ratio, vl / s);
        }
        if (ratio < T(1))
          for (uint32_t k = 0; k < ai[3]; k++)
            if (scalar_dof(ai[2] + k) && !fused_dof(ai[2] + k)) art[A_QD + ai[2] + k] *= ratio;
        bool any = false;
        for (uint32_t k = 0; k < ai[3]; k++) {
          uint32_t d = ai[2] + k;
          bool on = std::fabs(art[A_QD + d]) >= F_VLIM_FRAC * dof_passive(d).z;
          art[A_VLG + d] = on ? T(1) : T(0);
          any = any || (on && dof_passive(d).z >= T(0) && scalar_dof(d));
        }
        art[A_AVL + a] = any ? T(1) : T(0);
      }
      uint32_t n = 0;
      for (uint32_t k = 0; k < ai[3]; k++) {
        uint32_t d = ai[2] + k;
        V4<T> lim = dof_limit(d);
        bool on = false;
        if (lim.z > T(0.5)) {
          uint32_t l = tu(lim.w);
          T q = art[A_Q + link_qs(l) + d - link_dof(l)];
          T gg = F_JL_GAP + F_JL_SPEED * dt * std::fabs(art[A_QD + d]);
          on = q <= lim.x + gg || q >= lim.y - gg;
        }
        art[A_LG + d] = on ? T(1) : T(0);
        bool row = on || dof_passive(d).y > T(0) || JFL || (DRIVE_ROWS && driven_dof(d));
        if (row && scalar_dof(d)) { art[A_RL + ai[2] + n] = (T)d; n++; }
      }
      art[A_RLN + a] = (T)n;
      art_fk(a);
      art_factor(a);
    }
    if (VLIM_FREE) {
      // Free-body velocity limits: per axis, whether the body's speed is near its limit (wgsl.js:1878-1887).
      for (uint32_t b = 0; b < NB; b++) {
        if (body_kind((int)b) != 1) continue;
        V4<T> lm = body_limits((int)b);
        V3<T> v = xyz(vel[b * 2]), w = xyz(vel[b * 2 + 1]);
        for (uint32_t k = 0; k < 3; k++) {
          art[A_VLG + ND + b * 6 + k] = std::fabs(v[k]) >= F_VLIM_FRAC * lm.y ? T(1) : T(0);
          art[A_VLG + ND + b * 6 + 3 + k] = std::fabs(w[k]) >= F_VLIM_FRAC * lm.z ? T(1) : T(0);
        }
      }
    }
    // Phase 1
    uint32_t nc = std::min(counters[1], MAXC);
    uint32_t unit = 0, dn = 0;
    for (uint32_t c = 0; c < nc; c++) {
      const V4<T>& m = contacts[c].q[2];
      int a = ti(m.x), b = ti(m.y);
      uint64_t pk = contact_pair(c);
      uint32_t cls = 0xfffffffeu;  // no rows: both nodes asleep (or static)
      if (!pair_asleep(a, b)) {
        if (same_art(a, b)) {
          cls = 0xffffffffu;
          bool rws = !node_asleep(body_node(a));
          if (rws) dn++;
          S_CLOCAL[c] = rws ? 1u : 0u;
        } else {
          cls = (c == 0 || contact_pair(c - 1) != pk) ? 1u : 0u;
        }
      }
      S_CSLOT[c] = cls;
    }
    ndc = dn;
    uint32_t run0 = 0;
    for (uint32_t c = 0; c < nc; c++) {
      uint32_t cls = S_CSLOT[c];
      if (cls >= 0xfffffffeu) continue;
      const V4<T>& m = contacts[c].q[2];
      int a = ti(m.x), b = ti(m.y);
      if (cls == 1u) { unit++; run0 = c; }
      uint32_t h = unit - 1;
      if (h >= UH) { S_CSLOT[c] = 0xfffffffeu; continue; }
      if (c == run0) S_HKEY[h] = contact_pair(c);
      S_HCNT[h]++;
      S_CSLOT[c] = h;
      S_CLOCAL[c] = c - run0;
      if (a >= 0 && body_kind(a) == 2) wlinkhit[(uint32_t)a] = 1;
      if (b >= 0 && body_kind(b) == 2) wlinkhit[(uint32_t)b] = 1;
    }
    // Phase 2
    uint32_t u = 0, woff = 0;
    for (uint32_t h = 0; h < UH; h++) {
      uint64_t key = S_HKEY[h];
      if (key == 0) continue;
      if (u < MAXU) { wkey[u] = key; S_USIZE_TMP[u] = S_HCNT[h]; }
      S_SLOT2U[h] = u;
      S_UWOFF[h] = woff;
      woff += S_HCNT[h];
      u++;
    }
    wnu = std::min(u, MAXU);
    wnwalk = woff;
    for (uint32_t l = 0; l < NL; l++) {
      if (wlinkhit[(uint32_t)link_body(l)] == 0) continue;
      int j = (int)l;
      while (j >= 0) { wjsplit[(uint32_t)j]++; j = link_parent((uint32_t)j); }
    }
    for (uint32_t l = 0; l < NL; l++) {
      if (wlinkhit[(uint32_t)link_body(l)] == 0) continue;
      std::vector<uint32_t> path(MAXPATH), cnt(MAXPATH);
      uint32_t np = link_path(l, path.data());
      for (uint32_t i = 0; i < np; i++) cnt[i] = wjsplit[path[i]];
      link_response(l, cnt.data());
    }
    for (uint32_t d = 0; d < ND; d++) {
      V4<T> lim = dof_limit(d);
      if (lim.z > T(0.5) || dof_passive(d).y > T(0) || dof_passive(d).z >= T(0) || statF[SF_DOF + d * 5 + 3].w > T(0) ||
          (DRIVE_ROWS && driven_dof(d))) {
        uint32_t l = tu(lim.w);
        art[A_HINV + d] = dof_hinv(l, d - link_dof(l));
      }
    }
    for (uint32_t c = 0; c < nc; c++) {
      uint32_t h = S_CSLOT[c];
      if (h >= UH) continue;
      uint32_t u0 = S_SLOT2U[h];
      S_WALK[S_UWOFF[h] + S_CLOCAL[c]] = u0 >= MAXU ? 0xffffffffu : c;
    }
    uint32_t nu = wnu;
    S_META[0] = nu; S_META[2] = nc;
    for (uint32_t i = 0; i < nu; i++) { S_UKEY[i] = wkey[i]; S_USIZE[i] = S_USIZE_TMP[i]; wcol[i] = i; }
    uint32_t off = 0;
    for (uint32_t i = 0; i < nu; i++) { S_USTART[i] = off; off += S_USIZE[i]; }
    S_USTART[nu] = off;
    // Phase 5
    for (uint32_t c = 0; c < nc; c++) {
      uint32_t slot = S_CSLOT[c];
      if (slot >= 0xfffffffeu) continue;
      uint32_t u0 = S_SLOT2U[slot];
      if (u0 >= MAXU) continue;
      S_CORDER[S_USTART[wcol[u0]] + S_CLOCAL[c]] = c;
    }
    uint32_t nl = std::min(S_USTART[nu], MAXC);
    for (uint32_t p = 0; p < nl; p++) { BPC gg = bp_contact(S_CORDER[p]); S_RK0[p] = gg.k[0]; S_RK1[p] = gg.k[1]; S_RK2[p] = gg.k[2]; }
    for (uint32_t sp = 0; sp < nu; sp++) S_OSTART[sp] = S_USTART[sp];
    for (uint32_t p = 0; p < nl; p++) {
      uint32_t st = S_USTART[unit_of(p, nu)];
      uint32_t lead = p;
      for (uint32_t m = st; m < p; m++)
        if (S_RK2[m] == S_RK2[p] && S_RK0[m] == S_RK0[p] && S_RK1[m] == S_RK1[p]) { lead = m; break; }
      S_LEAD[p] = lead;
      for (uint32_t q = 0; q < 26; q++) S_SLOT[p * 26 + q] = 0;
    }
    for (uint32_t p = 0; p < nl; p++) {
      BPC gg = bp_contact(S_CORDER[p]);
      uint32_t sl = S_LEAD[p] * 26;
      for (uint32_t q = 0; q < 13; q++) { uint32_t v[2]; bp_value(gg, q, v); if (v[1] != 0) S_SLOT[sl + q * 2] = std::max(S_SLOT[sl + q * 2], v[0]); }
    }
    for (uint32_t p = 0; p < nl; p++) {
      BPC gg = bp_contact(S_CORDER[p]);
      uint32_t sl = S_LEAD[p] * 26;
      for (uint32_t q = 0; q < 13; q++) {
        uint32_t v[2]; bp_value(gg, q, v);
        if (v[1] != 0 && v[0] == S_SLOT[sl + q * 2]) S_SLOT[sl + q * 2 + 1] = std::max(S_SLOT[sl + q * 2 + 1], v[1]);
      }
    }
    for (uint32_t p = 0; p < nl; p++) {
      BPC gg = bp_contact(S_CORDER[p]);
      uint32_t sl = S_LEAD[p] * 26;
      bool win = false;
      for (uint32_t q = 0; q < 13; q++) {
        uint32_t v[2]; bp_value(gg, q, v);
        if (v[1] != 0 && v[0] == S_SLOT[sl + q * 2] && v[1] == S_SLOT[sl + q * 2 + 1]) win = true;
      }
      S_LEAD[p] = (win || COMPL) ? 1 : 0;  // compliant contacts stay whole
    }
    for (uint32_t sp = 0; sp < nu; sp++) {
      uint32_t n = S_USIZE[sp], st = S_USTART[sp], kept = 0;
      for (uint32_t j = 0; j < n; j++) {
        if (st + j >= MAXC) break;
        if (S_LEAD[st + j] != 0) { S_KEPT[st + kept] = S_CORDER[st + j]; kept++; }
      }
      S_USIZE[sp] = kept;
    }
    off = 0;
    for (uint32_t i = 0; i < nu; i++) { S_USTART[i] = off; off += S_USIZE[i]; }
    S_USTART[nu] = off;
    // Phase 6: rows
    uint32_t total = std::min(S_USTART[nu], MAXC);
    for (uint32_t idx = 0; idx < total; idx++) {
      uint32_t sp = unit_of(idx, nu);
      uint32_t c = S_KEPT[S_OSTART[sp] + idx - S_USTART[sp]];
      const Contact<T>& C = contacts[c];
      V4<T> c0 = C.q[0], c1 = C.q[1], cmeta = C.q[2];
      int a = ti(cmeta.x), b = ti(cmeta.y);
      uint32_t key = std::max(tu(cmeta.z) | (tu(cmeta.w) << 16), 1u);
      V3<T> p = xyz(c0);
      T phi = c0.w;
      V3<T> n = xyz(c1);
      T mu = phi <= F_FGAP ? c1.w * F_FSCALE : T(0);
      V3<T> t1 = cross(n, V3<T>{1, 0, 0});
      if (dot(t1, t1) < T(0.1)) t1 = cross(n, V3<T>{0, 1, 0});
      t1 = normalize(t1);
      V3<T> t2 = cross(n, t1);
      V3<T> ca{0, 0, 0}, cb{0, 0, 0};
      if (a >= 0) ca = xyz(poses[P_BODY + (uint32_t)a * 2]);
      if (b >= 0) cb = xyz(poses[P_BODY + (uint32_t)b * 2]);
      T half = T(0.5) * phi;
      V3<T> ra = p - n * half - ca, rb = p + n * half - cb;
      V3<T> fa = ra, fb = rb;
      uint32_t o = idx * ROWSN;
      V3<T> dirs[3] = {n, t1, t2};
      V3<T> Ma[3][2], Mb[3][2];
      for (uint32_t r = 0; r < 3; r++) {
        V3<T> d = dirs[r];
        V3<T> aa = r == 0 ? ra : fa, ab = r == 0 ? rb : fb;
        body_response(a, -d, -cross(aa, d), Ma[r]);
        body_response(b, d, cross(ab, d), Mb[r]);
        rows[o + 7 + r * 4] = v4(Ma[r][0], T(0)); rows[o + 8 + r * 4] = v4(Ma[r][1], T(0));
        rows[o + 9 + r * 4] = v4(Mb[r][0], T(0)); rows[o + 10 + r * 4] = v4(Mb[r][1], T(0));
      }
      T K[9];
      for (uint32_t r = 0; r < 3; r++)
        for (uint32_t s = 0; s < 3; s++) {
          V3<T> d = dirs[r];
          V3<T> aa = r == 0 ? ra : fa, ab = r == 0 ? rb : fb;
          K[r * 3 + s] = dot(-d, Ma[s][0]) + dot(-cross(aa, d), Ma[s][1]) + dot(d, Mb[s][0]) + dot(cross(ab, d), Mb[s][1]);
        }
      if (EXTRA) { rows[o + PO] = v4(fa, T(1)); rows[o + PO + 1] = v4(fb, T(0)); rows[o + PO + 2] = v4(p, T(0)); }
      if (TORS) { rows[o + TO + 3] = {0, 0, 0, C.q[3].w}; rows[o + TO + 4] = v4(p, T(0)); }
      if (FMODE) {
        rows[o + FMO] = {K[0], T(0.5) * (K[1] + K[3]), T(0.5) * (K[2] + K[6]), T(0)};
        rows[o + FMO + 1] = {K[4], T(0.5) * (K[5] + K[7]), K[8], T(0)};
      }
      T kscale = mx(K[0] + K[4] + K[8], T(1e-30));
      T effn = 0;
      if (K[0] > T(1e-6) * kscale && K[0] > T(1e-12)) effn = T(1) / (K[0] + F_CFM);
      T k00 = K[4] + T(1e-6) * kscale + F_CFM, k01 = T(0.5) * (K[5] + K[7]), k11 = K[8] + T(1e-6) * kscale + F_CFM;
      T det = k00 * k11 - k01 * k01;
      T rhs;
      if (phi < T(0)) rhs = mx(beta * phi / dt, -depen_cap(a, b)); else rhs = F_SPEC * phi / dt;
      T e = C.q[3].x;
      bool bounce = false;
      if (e > T(0)) {
        V3<T> ta0[2], tb0[2];
        pre_twist(a, ta0); pre_twist(b, tb0);
        T vn0 = dot(n, (tb0[0] + cross(tb0[1], rb)) - (ta0[0] + cross(ta0[1], ra)));
        bounce = restitution_fires(e, phi, vn0);
        if (bounce) rhs = e * vn0;
      }
      T ln = 0;
      V2<T> lt{0, 0};
      warm_lookup(key, mu, t1, t2, ln, lt);
      T effc = effn, rhsc = rhs;
      if (COMPL) {
        // contact_compliance: a spring-damper normal; the pair's ke and kd split over its kept points (wgsl.js:2194-2207)
        T kx = C.q[4].x / (T)std::max(S_USIZE[sp], 1u), ky = C.q[4].y / (T)std::max(S_USIZE[sp], 1u);
        T gam = 0;
        if (kx > T(0)) {
          T ce = phi <= T(0) ? ky : T(0);
          T den = dt * kx + ce;
          gam = T(1) / (dt * den);
          rhsc = kx * phi / den;
          if (effn > T(0)) effc = T(1) / (T(1) / effn + gam);
        }
        rows[o + CO] = {gam, 0, 0, 0};
      }
      rows[o + 0] = v4(n, effc);
      rows[o + 1] = v4(t1, rhsc);
      rows[o + 2] = v4(t2, mu);
      rows[o + 3] = v4(ra, (T)(key & 0xffffu));
      rows[o + 4] = v4(rb, (T)(key >> 16));
      T wrow = row_weight(phi);
      if (wrow < T(1) && ((a >= 0 && body_kind(a) == 2) || (b >= 0 && body_kind(b) == 2))) {
        V3<T> ua[2], ub[2];
        body_response_unsplit(a, -n, -cross(ra, n), ua);
        body_response_unsplit(b, n, cross(rb, n), ub);
        T du = dot(-n, ua[0]) + dot(-cross(ra, n), ua[1]) + dot(n, ub[0]) + dot(cross(rb, n), ub[1]);
        if (du > T(0) && K[0] != du) wrow = wrow * K[0] / (wrow * K[0] + (T(1) - wrow) * du);
      }
      rows[o + 5] = {k11 / det, -k01 / det, k00 / det, bounce ? T(1) : wrow};
      rows[o + 6] = {ln, lt.x, lt.y, bounce ? T(1e30) : phi};
      if (ANG) {
        // Spin (about n) and rolling (about t1, t2) rows: angular-only Jacobians (wgsl.js:2223-2243).
        V4<T> mat = C.q[3];
        V3<T> MA[3][2], MB[3][2];
        for (uint32_t r = 0; r < 3; r++) {
          V3<T> d = dirs[r];
          body_response(a, V3<T>{0, 0, 0}, -d, MA[r]);
          body_response(b, V3<T>{0, 0, 0}, d, MB[r]);
          rows[o + 22 + r * 4] = v4(MA[r][0], T(0)); rows[o + 23 + r * 4] = v4(MA[r][1], T(0));
          rows[o + 24 + r * 4] = v4(MB[r][0], T(0)); rows[o + 25 + r * 4] = v4(MB[r][1], T(0));
        }
        T KA[9];
        for (uint32_t r = 0; r < 3; r++)
          for (uint32_t s3 = 0; s3 < 3; s3++) KA[r * 3 + s3] = dot(-dirs[r], MA[s3][1]) + dot(dirs[r], MB[s3][1]);
        T ks = mx(KA[0], T(1e-12));
        T r00 = KA[4] + T(1e-9), r01 = T(0.5) * (KA[5] + KA[7]), r11 = KA[8] + T(1e-9);
        T rd = r00 * r11 - r01 * r01;
        rows[o + 19] = {mat.y, mat.z, KA[0] > T(1e-12) ? T(1) / ks : T(0), T(0)};
        rows[o + 20] = {r11 / rd, -r01 / rd, r00 / rd, T(0)};
        rows[o + 21] = {0, 0, 0, 0};
        if (ANG_ART) {
          // Raw 5x5 response H (t1, t2, spin, roll1, roll2), upper triangle, for the elliptic block (wgsl.js:2244-2267).
          V3<T> JA[5][2], JB[5][2], RA[5][2], RB[5][2];
          for (uint32_t q = 0; q < 2; q++) {
            V3<T> d = dirs[q + 1];
            JA[q][0] = -d; JA[q][1] = -cross(fa, d); JB[q][0] = d; JB[q][1] = cross(fb, d);
            RA[q][0] = Ma[q + 1][0]; RA[q][1] = Ma[q + 1][1]; RB[q][0] = Mb[q + 1][0]; RB[q][1] = Mb[q + 1][1];
          }
          for (uint32_t q = 0; q < 3; q++) {
            JA[q + 2][0] = {0, 0, 0}; JA[q + 2][1] = -dirs[q]; JB[q + 2][0] = {0, 0, 0}; JB[q + 2][1] = dirs[q];
            RA[q + 2][0] = MA[q][0]; RA[q + 2][1] = MA[q][1]; RB[q + 2][0] = MB[q][0]; RB[q + 2][1] = MB[q][1];
          }
          T Hu[16];
          uint32_t idx = 0;
          for (uint32_t x = 0; x < 5; x++)
            for (uint32_t y = x; y < 5; y++)
              Hu[idx++] = dot(JA[x][0], RA[y][0]) + dot(JA[x][1], RA[y][1]) + dot(JB[x][0], RB[y][0]) + dot(JB[x][1], RB[y][1]);
          Hu[15] = 0;
          for (uint32_t q = 0; q < 4; q++) rows[o + EO + q] = {Hu[q * 4], Hu[q * 4 + 1], Hu[q * 4 + 2], Hu[q * 4 + 3]};
        }
      }
    }
    if (TORS) for (uint32_t sp = 0; sp < nu; sp++) build_torsion(sp);
    if (PATCH) for (uint32_t sp = 0; sp < nu; sp++) build_patch(sp);
    // Phase 6c: coloring
    uint32_t totp = std::min(S_USTART[nu], MAXC);
    for (uint32_t c = 0; c < nc; c++) S_ROWOF[c] = 0xffffffffu;
    std::fill(wmask.begin(), wmask.end(), 0u);
    std::fill(wcount.begin(), wcount.end(), 0u);
    for (uint32_t sp = 0; sp < nu; sp++) { S_UTC[sp] = 0; wutp[sp] = 0; }
    for (uint32_t row = 0; row < totp; row++) {
      uint32_t sp = unit_of(row, nu);
      uint32_t c = S_KEPT[S_OSTART[sp] + row - S_USTART[sp]];
      S_RUNIT[row] = sp; S_RCON[row] = c;
      if (c < nc) S_ROWOF[c] = row;
    }
    uint32_t ncu = std::min(wnwalk, MAXC);
    for (uint32_t base = 0; base < ncu; base += MAXU) {
      uint32_t n = std::min(MAXU, ncu - base);
      for (uint32_t i = 0; i < n; i++) {
        uint32_t wc = S_WALK[base + i];
        uint32_t row = wc < MAXC ? S_ROWOF[std::min(wc, MAXC - 1)] : 0xffffffffu;
        wkey[i] = row < totp ? S_UKEY[S_RUNIT[std::min(row, MAXC - 1)]] : ~0ull;
      }
      for (uint32_t i = 0; i < n; i++) {
        uint64_t key = wkey[i];
        if (key != ~0ull) wcol[i] = first_free_color(key_a(key), key_b(key));
      }
      for (uint32_t i = 0; i < n; i++) {
        if (wkey[i] == ~0ull) continue;
        uint32_t c = S_WALK[base + i];
        uint32_t row = S_ROWOF[c];
        uint32_t col = wcol[i];
        S_PCOL[row] = col;
        wcount[col]++;
        uint64_t& utp = wutp[S_RUNIT[row]];
        utp = std::max(utp, ((uint64_t)(col + 1) << 32) | c);
      }
    }
    for (uint32_t sp = 0; sp < nu; sp++) {
      uint64_t packed = wutp[sp];
      S_UTC[sp] = (uint32_t)(packed >> 32);
      S_UTP[sp] = packed != 0 ? S_ROWOF[(uint32_t)packed] : 0xffffffffu;
    }
    uint32_t acc = 0, ncol = 0;
    for (uint32_t c = 0; c <= NCOL; c++) {
      uint32_t n = wcount[c];
      S_CSTART[c] = acc; wcount[c] = acc; acc += n;
      if (n > 0) ncol = c + 1;
    }
    S_CSTART[NCOL + 1] = acc;
    S_META[1] = ncol;
    for (uint32_t row = 0; row < totp; row++) {
      uint32_t slot = wcount[S_PCOL[row]]++;
      if (slot < MAXC) S_PORD[slot] = row;
    }
    // overflow color: by contact index (wgsl.js:1740-1780)
    {
      uint32_t o0 = std::min(S_CSTART[NCOL], MAXC);
      uint32_t m = std::min(S_CSTART[NCOL + 1], MAXC) - o0;
      if (m >= 2) {
        std::vector<uint32_t> rr(S_PORD.begin() + o0, S_PORD.begin() + o0 + m);
        std::sort(rr.begin(), rr.end(), [&](uint32_t x, uint32_t y) { return S_RCON[x] < S_RCON[y] || (S_RCON[x] == S_RCON[y] && x < y); });
        std::copy(rr.begin(), rr.end(), S_PORD.begin() + o0);
      }
    }
    // Phase 7: dense rows for same-articulation contacts, in contact order, then the mimics in fixed slots
    S_META[3] = ndc + loop_rows() + NMIMIC;
    uint32_t dslot = 0;
    for (uint32_t c = 0; c < nc; c++) {
      if (S_CSLOT[c] != 0xffffffffu || S_CLOCAL[c] == 0) continue;
      uint32_t slot = dslot++;
      if (slot >= MAXDENSE) continue;
      const Contact<T>& C = contacts[c];
      V4<T> c0 = C.q[0], c1 = C.q[1], cmeta = C.q[2];
      int a = ti(cmeta.x), b = ti(cmeta.y);
      uint32_t key = std::max(tu(cmeta.z) | (tu(cmeta.w) << 16), 1u);
      V3<T> p = xyz(c0);
      T phi = c0.w;
      V3<T> n = xyz(c1);
      T mu = phi <= F_FGAP ? c1.w * F_FSCALE : T(0);
      V3<T> t1 = cross(n, V3<T>{1, 0, 0});
      if (dot(t1, t1) < T(0.1)) t1 = cross(n, V3<T>{0, 1, 0});
      t1 = normalize(t1);
      V3<T> t2 = cross(n, t1);
      uint32_t art_id = (uint32_t)body_art(a);
      uint32_t ai[4];
      art_info(art_id, ai);
      std::vector<T> J(3 * NDA, T(0));
      add_gen_jac((uint32_t)body_link(b), p, T(1), n, t1, t2, ai[2], J.data());
      add_gen_jac((uint32_t)body_link(a), p, T(-1), n, t1, t2, ai[2], J.data());
      uint32_t o = A_DENSE + slot * DPS;
      T K[9];
      for (uint32_t r = 0; r < 3; r++) {
        std::vector<T> tau(NDA), y(NDA, T(0));
        for (uint32_t k = 0; k < NDA; k++) tau[k] = k < ai[3] ? J[r * NDA + k] : T(0);
        art_solve_gen(art_id, tau.data(), y.data());
        for (uint32_t k = 0; k < ai[3]; k++) { art[o + DHDR + r * NDA + k] = J[r * NDA + k]; art[o + DHDR + 3 * NDA + r * NDA + k] = y[k]; }
        for (uint32_t s2 = 0; s2 < 3; s2++) { T acc2 = 0; for (uint32_t k = 0; k < ai[3]; k++) acc2 += J[s2 * NDA + k] * y[k]; K[s2 * 3 + r] = acc2; }
      }
      T kscale = mx(K[0] + K[4] + K[8], T(1e-30));
      T effn = 0;
      if (K[0] > T(1e-6) * kscale && K[0] > T(1e-12)) effn = T(1) / (K[0] + F_CFM);
      T k00 = K[4] + T(1e-6) * kscale + F_CFM, k01 = T(0.5) * (K[5] + K[7]), k11 = K[8] + T(1e-6) * kscale + F_CFM;
      T det = k00 * k11 - k01 * k01;
      T rhs;
      if (phi < T(0)) rhs = mx(beta * phi / dt, -depen_cap(a, b)); else rhs = F_SPEC * phi / dt;
      T e = C.q[3].x;
      bool bounce = false;
      if (e > T(0)) {
        T vn0 = 0;
        for (uint32_t k = 0; k < ai[3]; k++) vn0 += J[k] * art[A_QD + ai[2] + k];
        bounce = restitution_fires(e, phi, vn0);
        if (bounce) rhs = e * vn0;
      }
      T ln = 0;
      V2<T> lt{0, 0};
      warm_lookup(key, mu, t1, t2, ln, lt);
      art[o] = (T)art_id; art[o + 1] = (T)(key & 0xffffu); art[o + 2] = (T)(key >> 16); art[o + 3] = mu;
      art[o + 4] = rhs; art[o + 5] = effn;
      art[o + 6] = k11 / det; art[o + 7] = -k01 / det; art[o + 8] = k00 / det;
      art[o + 9] = ln; art[o + 10] = lt.x; art[o + 11] = lt.y;
      st3(o + 12, t1); st3(o + 15, t2); st3(o + 18, n);
      art[o + 21] = 0;
      art[o + 24] = bounce ? T(1) : row_weight(phi);
      art[o + 22] = bounce ? T(1e30) : phi;
    }
    // Loop closures (Steel's connect rows, wgsl.js:2445-2480): three point rows at the anchors' midpoint, plus two
    // hinge-alignment rows for a revolute closure; a world parent contributes no dofs.
    uint32_t slot = ndc;
    for (uint32_t i = 0; i < NLOOP; i++) {
      V4<T> L0 = statF[SF_LOOP + i * 5];
      int pl = ti(L0.x);
      uint32_t cl = tu(L0.y), art_id = tu(L0.z);
      bool rev = L0.w > T(0.5);
      uint32_t s = slot;
      slot += rev ? 2 : 1;
      if (s + (rev ? 1 : 0) >= MAXDENSE) continue;
      V3<T> pa = xyz(statF[SF_LOOP + i * 5 + 1]);
      V4<T> rp{0, 0, 0, 1};
      if (pl >= 0) { rp = link_rot((uint32_t)pl); pa = ld3(lo((uint32_t)pl) + L_POS) + qrot(rp, pa); }
      V4<T> rc = link_rot(cl);
      V3<T> pc = ld3(lo(cl) + L_POS) + qrot(rc, xyz(statF[SF_LOOP + i * 5 + 2]));
      V3<T> pm = T(0.5) * (pa + pc);
      uint32_t ai[4];
      art_info(art_id, ai);
      V3<T> ex{1, 0, 0}, ey{0, 1, 0}, ez{0, 0, 1};
      std::vector<T> J(3 * NDA, T(0));
      add_gen_jac(cl, pm, T(1), ex, ey, ez, ai[2], J.data());
      if (pl >= 0) add_gen_jac((uint32_t)pl, pm, T(-1), ex, ey, ez, ai[2], J.data());
      put_bilateral(A_DENSE + s * DPS, art_id, J, 3, beta * (pc - pa) / dt);
      if (rev) {
        V3<T> ap = qrot(rp, xyz(statF[SF_LOOP + i * 5 + 3])), ac = qrot(rc, xyz(statF[SF_LOOP + i * 5 + 4]));
        V3<T> e1 = cross(ap, ex);
        if (dot(e1, e1) < T(1e-6)) e1 = cross(ap, ey);
        e1 = normalize(e1);
        V3<T> e2 = cross(ap, e1);
        std::vector<T> Ja(3 * NDA, T(0));
        add_gen_ang(cl, T(1), e1, e2, ai[2], Ja.data());
        if (pl >= 0) add_gen_ang((uint32_t)pl, T(-1), e1, e2, ai[2], Ja.data());
        V3<T> err = cross(ap, ac);
        put_bilateral(A_DENSE + (s + 1) * DPS, art_id, Ja, 2, beta * V3<T>{dot(e1, err), dot(e2, err), T(0)} / dt);
      }
    }
    // Mimic rows: qd_follower - c1 qd_leader, Baumgarte on q_follower - c1 q_leader - c0 (wgsl.js:2482-2492).
    for (uint32_t i = 0; i < NMIMIC; i++) {
      V4<T> M0 = statF[SF_MIMIC + i * 2], M1 = statF[SF_MIMIC + i * 2 + 1];
      uint32_t slot = ndc + loop_rows() + i;
      if (slot >= MAXDENSE) continue;
      std::vector<T> J(3 * NDA, T(0));
      J[tu(M0.y)] = T(1);
      J[tu(M0.z)] -= M0.w;
      T phi = art[A_Q + tu(M1.x)] - M0.w * art[A_Q + tu(M1.y)] - M1.z;
      put_bilateral(A_DENSE + slot * DPS, tu(M0.x), J, 1, V3<T>{beta * phi / dt, T(0), T(0)});
    }
  }
  // Bilateral dense block: nb unclamped rows of J, their Y = H^-1 J^T, 1 / (J Y) and the rhs (wgsl.js:1351-1369).
  void put_bilateral(uint32_t o, uint32_t art_id, const std::vector<T>& J, uint32_t nb, V3<T> rhs) {
    uint32_t ai[4];
    art_info(art_id, ai);
    for (uint32_t r = 0; r < nb; r++) {
      std::vector<T> tau(NDA), y(NDA, T(0));
      for (uint32_t k = 0; k < NDA; k++) tau[k] = k < ai[3] ? J[r * NDA + k] : T(0);
      art_solve_gen(art_id, tau.data(), y.data());
      T kk = 0;
      for (uint32_t k = 0; k < ai[3]; k++) {
        art[o + DHDR + r * NDA + k] = J[r * NDA + k];
        art[o + DHDR + 3 * NDA + r * NDA + k] = y[k];
        kk += J[r * NDA + k] * y[k];
      }
      art[o + 5 + r] = kk > T(1e-12) ? T(1) / kk : T(0);
    }
    art[o] = (T)art_id; art[o + 1] = 0; art[o + 2] = 0;
    art[o + 4] = rhs.x; art[o + 22] = rhs.y; art[o + 23] = rhs.z;
    art[o + 9] = 0; art[o + 10] = 0; art[o + 11] = 0;
    art[o + 21] = (T)nb;
  }

  // ============================ solve (wgsl.js:2501-3891) ============================
  // ---- Nesterov momentum on the colored contact rows (wgsl.js:3492-3557): at the start of each position sweep
  // the impulses move by beta along the previous sweep's change; the represented change e = y - lam reaches the
  // velocities through the row's own responses.  rows[o + MO] keeps lam at the last extrapolation, rows[o + MO + 1]
  // keeps e.
  void momentum_unit(uint32_t row, int a, int b, T beta, bool init) {
    uint32_t o = row * ROWS;
    V4<T> lam = rows[o + 6];
    V3<T> l3 = xyz(lam);
    if (init) { rows[o + MO] = v4(l3, T(0)); rows[o + MO + 1] = {0, 0, 0, 0}; return; }
    V3<T> y = l3 + beta * (l3 - xyz(rows[o + MO]));
    V3<T> e = y - l3;
    rows[o + MO] = v4(l3, T(0)); rows[o + MO + 1] = v4(e, T(0));
    if (e.x == T(0) && e.y == T(0) && e.z == T(0)) return;
    rows[o + 6] = v4(y, lam.w);
    V3<T> ta[2], tb[2];
    load_tw(a, ta); load_tw(b, tb);
    ta[0] += xyz(rows[o + 7]) * e.x + xyz(rows[o + 11]) * e.y + xyz(rows[o + 15]) * e.z;
    ta[1] += xyz(rows[o + 8]) * e.x + xyz(rows[o + 12]) * e.y + xyz(rows[o + 16]) * e.z;
    tb[0] += xyz(rows[o + 9]) * e.x + xyz(rows[o + 13]) * e.y + xyz(rows[o + 17]) * e.z;
    tb[1] += xyz(rows[o + 10]) * e.x + xyz(rows[o + 14]) * e.y + xyz(rows[o + 18]) * e.z;
    V3<T> ra = xyz(rows[o + 3]), rb = xyz(rows[o + 4]), fa = ra, fb = rb;
    if (EXTRA) { fa = xyz(rows[o + PO]); fb = xyz(rows[o + PO + 1]); }
    V3<T> fn = xyz(rows[o]) * e.x, ft = xyz(rows[o + 1]) * e.y + xyz(rows[o + 2]) * e.z;
    add_imp(a, -fn - ft, -cross(ra, fn) - cross(fa, ft));
    add_imp(b, fn + ft, cross(rb, fn) + cross(fb, ft));
    if (a >= 0) { tw[(uint32_t)a * 2] = v4(ta[0], T(0)); tw[(uint32_t)a * 2 + 1] = v4(ta[1], T(0)); }
    if (b >= 0) { tw[(uint32_t)b * 2] = v4(tb[0], T(0)); tw[(uint32_t)b * 2 + 1] = v4(tb[1], T(0)); }
  }
  // One momentum step before position sweep `it`: the O'Donoghue-Candes restart (k = 0) when the last sweep's
  // correction opposed its extrapolation (sum over rows of (lam - lam_last - e) . e below zero), else k + 1 and
  // beta = (k - 1) / (k + 2) for k > 1.  The sum runs as 256 strided partials (partial j over rows j, j + 256, ...)
  // added in order, the reference's workgroup reduction.
  void momentum_step(uint32_t ncol, uint32_t it, const std::vector<uint32_t>& wcs, const std::vector<uint32_t>& slot) {
    uint32_t nrows = std::min(wcs[NCOL + 1], MAXC);
    if (it > 0) {
      T part[256];
      for (uint32_t j = 0; j < 256; j++) part[j] = 0;
      for (uint32_t u = 0; u < nrows; u++) {
        uint32_t o = S_PORD[u] * ROWS;
        V3<T> e = xyz(rows[o + MO + 1]);
        part[u % 256] += dot(xyz(rows[o + 6]) - xyz(rows[o + MO]) - e, e);
      }
      T total = 0;
      for (uint32_t j = 0; j < 256; j++) total += part[j];
      mom_k = (total < T(0)) ? T(0) : mom_k + T(1);
      mom_beta = mom_k > T(1) ? (mom_k - T(1)) / (mom_k + T(2)) : T(0);
    } else {
      mom_k = 0; mom_beta = 0;
    }
    for (uint32_t c = 0; c < NCOL && c < ncol; c++)
      for (uint32_t u = wcs[c]; u < wcs[c + 1]; u++) momentum_unit(slot[u * 4], (int)slot[u * 4 + 2], (int)slot[u * 4 + 3], mom_beta, it == 0);
    for (uint32_t u = wcs[NCOL]; u < wcs[NCOL + 1]; u++) {
      uint32_t uu = std::min(u, MAXC - 1);
      momentum_unit(slot[uu * 4], (int)slot[uu * 4 + 2], (int)slot[uu * 4 + 3], mom_beta, it == 0);
    }
    for (uint32_t a = 0; a < NA; a++) {
      art_contact_impulses(a);
      art_velocity_limits(a);
    }
  }
  // Twists of the velocity pass's snapshot (wgsl.js:2930-2934).
  void snap_tw(int b, V3<T> out[2]) const {
    if (b < 0) { out[0] = {0, 0, 0}; out[1] = {0, 0, 0}; return; }
    out[0] = ld3(A_SNAP + (uint32_t)b * 8); out[1] = ld3(A_SNAP + (uint32_t)b * 8 + 4);
  }
  // Velocity pass (velocity_iterations, wgsl.js:3700-3736): positions integrate with the velocities reached by the
  // position sweeps (kept as a snapshot); bias-free sweeps from the accumulated impulses then set the output
  // velocities.  Contact right-hand sides become 0, or a speculative gap still open at the snapshot velocity.
  void velocity_pass(uint32_t ncol, const std::vector<uint32_t>& wcs, const std::vector<uint32_t>& slot, uint32_t nu) {
    for (uint32_t b = 0; b < NB; b++) { st3(A_SNAP + b * 8, xyz(tw[b * 2])); st3(A_SNAP + b * 8 + 4, xyz(tw[b * 2 + 1])); }
    for (uint32_t i = 0; i < ND; i++) art[A_SNAP + NB * 8 + i] = aqd[i];
    for (uint32_t us = 0; us < nu; us++) {
      uint64_t key = S_UKEY[us];
      int a = key_a(key), b = key_b(key);
      V3<T> ta[2], tb[2];
      load_tw(a, ta); load_tw(b, tb);
      uint32_t s0 = S_USTART[us];
      for (uint32_t i = s0; i < std::min(s0 + S_USIZE[us], MAXC); i++) {
        uint32_t o = i * ROWS;
        T phi = rows[o + 6].w;
        if (phi >= T(1e29)) continue;
        V3<T> n = xyz(rows[o]);
        T jv = dot(n, (tb[0] + cross(tb[1], xyz(rows[o + 4]))) - (ta[0] + cross(ta[1], xyz(rows[o + 3]))));
        rows[o + 1].w = (phi > T(0) && phi + dt * jv > T(1e-6)) ? phi / dt : T(0);
      }
    }
    uint32_t nd = std::min(S_META[3], MAXDENSE);
    for (uint32_t i = 0; i < nd; i++) {
      uint32_t o = A_DENSE + i * DPS;
      if (art[o + 21] > T(0.5) || art[o + 22] >= T(1e29)) continue;
      uint32_t ai[4];
      art_info(tu(art[o]), ai);
      T jv = 0;
      for (uint32_t k = 0; k < ai[3]; k++) jv += art[o + DHDR + k] * aqd[ai[2] + k];
      T phi = art[o + 22];
      art[o + 4] = (phi > T(0) && phi + dt * jv > T(1e-6)) ? phi / dt : T(0);
    }
    vpass = true;
    for (uint32_t it = 0; it < VELITERS; it++) sweep(1, ncol, ITERS + it, wcs, slot, nu);
    vpass = false;
  }
  void load_tw(int b, V3<T> out[2]) const {
    if (b < 0) { out[0] = {0, 0, 0}; out[1] = {0, 0, 0}; return; }
    out[0] = xyz(tw[(uint32_t)b * 2]); out[1] = xyz(tw[(uint32_t)b * 2 + 1]);
  }
  void add_imp(int b, V3<T> f, V3<T> t) {
    if (b < 0) return;
    if (body_kind(b) != 2) return;
    imp[(uint32_t)b * 2] = imp[(uint32_t)b * 2] + v4(f, T(0));
    imp[(uint32_t)b * 2 + 1] = imp[(uint32_t)b * 2 + 1] + v4(t, T(0));
  }
  // ---- Steel's joint torsional/rolling cone for articulated contacts (PropagationAngularFrictionBlock, elliptic cone;
  // wgsl.js:2655-2925): rows (t1, t2, spin, roll1, roll2) minimize their quadratic over the ball |lambda / mu| <= N,
  // by a More-Sorensen trust region (proximal fallback), relaxed by omega and line-searched from the current
  // impulses moved into the cone.
  struct Ball { T y[5]; bool ok, interior; };
  static T cone_norm(const T* y) { T s = 0; for (uint32_t k = 0; k < 5; k++) s += y[k] * y[k]; return std::sqrt(s); }
  static void cone_project(T* y, T load) {
    T n = cone_norm(y);
    if (n > load) for (uint32_t k = 0; k < 5; k++) y[k] *= load / n;
  }
  static bool live5(const T* H, const bool* act, uint32_t k) { return act[k] && H[k * 5 + k] > T(0); }
  // Cholesky of H + alpha I over the live rows; fails when a pivot keeps under 1e-6 of its diagonal.
  static bool factor5(const T* H, const bool* act, T alpha, T* L) {
    for (uint32_t k = 0; k < 5; k++)
      for (uint32_t l = 0; l <= k; l++) {
        T v = (live5(H, act, k) && live5(H, act, l)) ? H[k * 5 + l] : T(0);
        if (k == l) v = live5(H, act, k) ? v + alpha : T(1);
        for (uint32_t m = 0; m < l; m++) v -= L[k * 5 + m] * L[l * 5 + m];
        if (k == l) {
          if (!(v > T(1e-6) * (live5(H, act, k) ? H[k * 5 + k] + alpha : T(1)))) return false;
          L[k * 5 + k] = std::sqrt(v);
        } else {
          L[k * 5 + l] = v / L[l * 5 + l];
        }
      }
    return true;
  }
  static void lower5(const T* L, T* x) {
    for (uint32_t k = 0; k < 5; k++) {
      for (uint32_t m = 0; m < k; m++) x[k] -= L[k * 5 + m] * x[m];
      x[k] /= L[k * 5 + k];
    }
  }
  static void upper5(const T* L, T* x) {
    for (uint32_t kk = 0; kk < 5; kk++) {
      uint32_t k = 4 - kk;
      for (uint32_t m = k + 1; m < 5; m++) x[k] -= L[m * 5 + k] * x[m];
      x[k] /= L[k * 5 + k];
    }
  }
  // y = -(H + alpha I)^-1 c with alpha by Newton on 1/|y(alpha)| inside a bracket; ok when it met the 1e-5 boundary
  // residual or found an interior minimizer.
  static Ball cholesky_ball(const T* H, const T* c, const bool* act, T load) {
    Ball r;
    T b[5], cn = 0;
    for (uint32_t k = 0; k < 5; k++) { b[k] = live5(H, act, k) ? c[k] : T(0); cn += b[k] * b[k]; r.y[k] = 0; }
    cn = std::sqrt(cn);
    r.interior = !(cn > T(0));
    r.ok = true;
    if (!(load > T(0) && cn > T(0))) return r;
    r.ok = false;
    T lo = 0, hi = cn / load, alpha = 0;
    for (uint32_t it = 0; it < 40; it++) {
      T nx = -1;
      T L[25] = {};
      if (factor5(H, act, alpha, L)) {
        T yy[5];
        for (uint32_t k = 0; k < 5; k++) yy[k] = -b[k];
        lower5(L, yy); upper5(L, yy);
        for (uint32_t k = 0; k < 5; k++) r.y[k] = yy[k];
        T n = cone_norm(r.y);
        if (alpha == T(0) && n <= load) { r.interior = true; r.ok = true; return r; }
        if (std::fabs(n - load) <= T(1e-5) * load) { r.ok = true; break; }
        if (n < load) hi = alpha; else lo = alpha;
        T wv[5];
        for (uint32_t k = 0; k < 5; k++) wv[k] = r.y[k];
        lower5(L, wv);
        T wn = 0;
        for (uint32_t k = 0; k < 5; k++) wn += wv[k] * wv[k];
        nx = alpha + n * n / wn * (n - load) / load;
      } else {
        lo = alpha;
      }
      if (hi - lo <= T(1e-4) * hi) break;
      alpha = (nx > lo && nx < hi) ? nx : (lo > T(0) ? std::sqrt(lo * hi) : T(1e-3) * hi);
    }
    T n = cone_norm(r.y);
    if (n > load) for (uint32_t k = 0; k < 5; k++) r.y[k] *= load / n;
    return r;
  }
  // Minimize 0.5 y'Hy + c'y over |y| <= load; when dependent rows defeat Cholesky, a proximal term
  // 0.5 rho (y - center/scale)' diag(H) (y - center/scale), rho raised tenfold until the ball solve converges.
  static Ball trust_region(const T* H, const T* c, const bool* act, T load, const T* center, const T* scale) {
    Ball first = cholesky_ball(H, c, act, load);
    if (first.ok) return first;
    T ce[5];
    for (uint32_t k = 0; k < 5; k++) ce[k] = (act[k] && scale[k] > T(0)) ? center[k] / scale[k] : T(0);
    T rho = T(1e-5);
    for (uint32_t it = 0; it < 7; it++) {
      T K[25], g[5];
      for (uint32_t i = 0; i < 25; i++) K[i] = H[i];
      for (uint32_t k = 0; k < 5; k++) {
        T wt = live5(H, act, k) ? rho * H[k * 5 + k] : T(0);
        K[k * 5 + k] += wt;
        g[k] = c[k] - wt * ce[k];
      }
      rho *= T(10);
      Ball t = cholesky_ball(K, g, act, load);
      if (!t.ok) continue;
      bool fin = true;
      for (uint32_t k = 0; k < 5; k++) fin = fin && std::fabs(t.y[k]) < T(3.0e38);
      if (fin) return t;
    }
    return first;
  }
  // One block visit: H0, g0 the rows' response and current velocity residual, mu the row coefficients, lam the
  // current impulses, load the normal impulse; returns the new impulses.
  void ell_block(const T* H0, const T* g0, const T* mu_in, const T* lam, T load_in, T* out) const {
    T load = mx(load_in, T(0));
    T mu[5], H[25], gr[5], big = 0;
    for (uint32_t i = 0; i < 25; i++) H[i] = H0[i];
    for (uint32_t k = 0; k < 5; k++) { gr[k] = g0[k]; mu[k] = mx(mu_in[k], T(0)); big = mx(big, H[k * 5 + k]); }
    bool act[5];
    T comp[5];
    for (uint32_t k = 0; k < 5; k++) {
      act[k] = mu[k] > T(0) && H[k * 5 + k] > T(1e-12) * big;
      comp[k] = (k >= 2 && act[k] && load > T(0)) ? F_CREEP / (mu[k] * mu[k] * load) : T(0);
      H[k * 5 + k] += comp[k]; gr[k] += comp[k] * lam[k];
    }
    T y0[5], c[5], y[5];
    for (uint32_t k = 0; k < 5; k++) {
      y0[k] = act[k] ? lam[k] / mu[k] : T(0);
      gr[k] = act[k] ? gr[k] * mu[k] : T(0);
      for (uint32_t l = 0; l < 5; l++) H[k * 5 + l] *= mu[k] * mu[l];
    }
    for (uint32_t k = 0; k < 5; k++) {
      T v = gr[k];
      for (uint32_t l = 0; l < 5; l++) v -= H[k * 5 + l] * y0[l];
      c[k] = act[k] ? v : T(0);
    }
    Ball tr = trust_region(H, c, act, load, lam, mu_in);
    for (uint32_t k = 0; k < 5; k++) y[k] = tr.y[k];
    bool stick = tr.interior && cone_norm(y) <= load;
    if (!stick) cone_project(y, load);
    T stp[5], shift[5];
    for (uint32_t k = 0; k < 5; k++) stp[k] = lam[k] + F_OMEGA * (mu[k] * y[k] - lam[k]);
    if (F_OMEGA != T(1)) {
      for (uint32_t k = 0; k < 5; k++) y[k] = act[k] ? stp[k] / mu_in[k] : T(0);
      cone_project(y, load);
      for (uint32_t k = 0; k < 5; k++) stp[k] = mu[k] * y[k];
    }
    for (uint32_t k = 0; k < 5; k++) y[k] = act[k] ? lam[k] / mu_in[k] : T(0);
    cone_project(y, load);
    for (uint32_t k = 0; k < 5; k++) { T b = mu[k] * y[k]; shift[k] = b - lam[k]; stp[k] -= b; }
    T slope = 0, curv = 0;
    for (uint32_t k = 0; k < 5; k++) {
      T hs = 0, hp = 0;
      for (uint32_t l = 0; l < 5; l++) { hs += H0[k * 5 + l] * shift[l]; hp += H0[k * 5 + l] * stp[l]; }
      slope += stp[k] * (g0[k] + hs + comp[k] * (lam[k] + shift[k]));
      curv += stp[k] * (hp + comp[k] * stp[k]);
    }
    T t = slope < T(0) ? (curv > -slope ? -slope / curv : T(1)) : T(0);
    for (uint32_t k = 0; k < 5; k++) out[k] = lam[k] + shift[k] + (t > T(0) ? t * stp[k] : T(0));
  }
  // ---- Steel's coupled contact solves (friction_mode bisection, bisection_desaxce, coulomb_newton; wgsl.js:2564-2653):
  // u = live (normal, t1, t2) velocities, G the 3x3 Delassus block (nn, nt1, nt2 | t1t1, t1t2, t2t2).
  static V2<T> tangent_step(V2<T> ut, V4<T> g1) {
    T det = g1.x * g1.z - g1.y * g1.y;
    if (std::fabs(det) <= T(1e-20)) return {0, 0};
    return {(-ut.x * g1.z + ut.y * g1.y) / det, (ut.x * g1.y - ut.y * g1.x) / det};
  }
  static V2<T> clip_cone(V2<T> t, T lim) {
    T m = length(t);
    if (m > lim && m > T(1e-20)) return t * (lim / m);
    return t;
  }
  static V3<T> fixed_normal_pair(V2<T> ut0, V4<T> g1, T mu, V3<T> lam) {
    V2<T> f = friction_pair_solve(g1.x, g1.y, g1.z, ut0.x, ut0.y, lam.y, lam.z, mx(mu * lam.x, T(0)), T(1));
    return {lam.x, f.x, f.y};
  }
  static V3<T> contact_block(uint32_t mode, T un, V2<T> ut0, V4<T> g0, V4<T> g1, T vtarget0, T mu, V3<T> lam) {
    T gnn = g0.x;
    V2<T> gnt{g0.y, g0.z};
    T trace = g1.x + g1.z;
    if (g1.x * g1.z - g1.y * g1.y <= T(1e-6) * trace * trace) return fixed_normal_pair(ut0, g1, mu, lam);
    if (gnn < T(1e-20)) return lam;
    V2<T> lt{lam.y, lam.z};
    T vt_n = vtarget0;
    if (mode == 2) vt_n += mu * length(ut0);
    if (mode == 3) {
      T un_f = un - (gnn * lam.x + dot(gnt, lt));
      V2<T> ut_f = ut0 - V2<T>{gnt.x * lam.x + g1.x * lt.x + g1.y * lt.y, gnt.y * lam.x + g1.y * lt.x + g1.z * lt.y};
      if (un_f >= vt_n) return {0, 0, 0};
      T bN = un_f - vt_n;
      T A00 = g1.x - gnt.x * gnt.x / gnn, A01 = g1.y - gnt.x * gnt.y / gnn, A11 = g1.z - gnt.y * gnt.y / gnn;
      if (A00 * A11 - A01 * A01 <= T(1e-6) * trace * trace) return fixed_normal_pair(ut0, g1, mu, lam);
      V2<T> c = ut_f - (bN / gnn) * gnt;
      T id0 = T(1) / (A00 * A11 - A01 * A01);
      V2<T> s = V2<T>{A11 * c.x - A01 * c.y, A00 * c.y - A01 * c.x} * id0;
      T phi0 = length(s) - mu * (dot(gnt, s) - bN) / gnn;
      if (phi0 > T(0)) {
        T hi = 1;
        for (uint32_t e = 0; e < 30; e++) {
          T a = A00 + hi, d = A11 + hi, idt = T(1) / (a * d - A01 * A01);
          V2<T> sh = V2<T>{d * c.x - A01 * c.y, a * c.y - A01 * c.x} * idt;
          if (length(sh) - mu * (dot(gnt, sh) - bN) / gnn < T(0)) break;
          hi *= T(2);
        }
        T lo = 0, x = T(0.5) * hi;
        T tol = T(1e-6) + T(1e-6) * phi0;
        for (uint32_t k = 0; k < 50; k++) {
          T a = A00 + x, d = A11 + x, idt = T(1) / (a * d - A01 * A01);
          V2<T> sx = V2<T>{d * c.x - A01 * c.y, a * c.y - A01 * c.x} * idt;
          V2<T> tx = V2<T>{d * sx.x - A01 * sx.y, a * sx.y - A01 * sx.x} * idt;
          T ns = length(sx);
          T fx = ns - mu * (dot(gnt, sx) - bN) / gnn;
          T dfx = -dot(sx, tx) / ns + mu * dot(gnt, tx) / gnn;
          s = sx;
          if (std::fabs(fx) < tol || std::fabs(hi - lo) < T(1e-6) * (T(1) + hi)) break;
          if (fx > T(0)) lo = x; else hi = x;
          T xn = T(0.5) * (lo + hi);
          if (dfx != T(0)) { T xt = x - fx / dfx; if (xt > lo && xt < hi) xn = xt; }
          x = xn;
        }
      }
      V2<T> rt = T(-1) * s;
      T ln = mx(-(dot(gnt, rt) + bN) / gnn, T(0));
      T lim = mx(mu * ln, T(0));
      T m = length(rt);
      V2<T> r2 = m > lim ? rt * (lim / m) : rt;
      return {ln, r2.x, r2.y};
    }
    // Bisection on the normal impulse with the tangent pair re-solved and clipped at each trial.
    T u0 = un - gnn * lam.x - dot(gnt, lt);
    if (u0 >= vt_n) return {0, 0, 0};
    T lo = 0;
    T hi = mx(mx(T(2) * lam.x, (vt_n - un) / gnn + lam.x), T(1));
    T flo = u0 - vt_n, fhi = -1;
    for (uint32_t k = 0; k < 20; k++) {
      T mid = T(0.5) * (lo + hi);
      T dn = mid - lam.x;
      V2<T> tr = clip_cone(lt + tangent_step(ut0 + gnt * dn, g1), mu * mid);
      T u = un + gnn * dn + dot(gnt, tr - lt);
      if (u < vt_n) { lo = mid; flo = u - vt_n; } else { hi = mid; fhi = u - vt_n; }
    }
    T ln = T(0.5) * (lo + hi);
    if (fhi >= T(0)) ln = lo - flo * (hi - lo) / (fhi - flo);
    V2<T> tf = clip_cone(lt + tangent_step(ut0 + gnt * (ln - lam.x), g1), mu * ln);
    return {ln, tf.x, tf.y};
  }
  static V2<T> friction_pair_solve(T a0, T c0, T d0, T r0, T r1, T old0, T old1, T radius, T omega) {  // wgsl.js:2952-3011
    if (radius <= T(0)) return {0, 0};
    T scale = mx(mx(a0, d0), T(1e-20));
    T a = a0 / scale, c = c0 / scale, d = d0 / scale;
    T largest = T(0.5) * (a + d + std::sqrt((a - d) * (a - d) + T(4) * c * c));
    T smallest = largest > T(0) ? mx((a * d - c * c) / largest, T(0)) : T(0);
    T vx = c, vy = largest - a;
    if (a >= d) { vx = largest - d; vy = c; }
    T vn = std::sqrt(vx * vx + vy * vy);
    if (vn > T(0)) { vx /= vn; vy /= vn; } else { vx = 1; vy = 0; }
    T r0r = (vx * r0 + vy * r1) / scale;
    T r1r = (-vy * r0 + vx * r1) / scale;
    V2<T> res{old0, old1};
    bool sticking = false;
    if (largest > T(0) && (smallest > T(0) || r1r == T(0))) {
      T dx = r0r / largest;
      T dy = smallest > T(0) ? r1r / smallest : T(0);
      res = {old0 - (vx * dx - vy * dy), old1 - (vy * dx + vx * dy)};
      sticking = length(res) <= radius;
    }
    if (!sticking) {
      T on = length(V2<T>{old0, old1}), rn = length(V2<T>{r0, r1});
      if (on > T(0) && on <= radius && radius - on <= T(2e-7) * radius && r0 * old0 + r1 * old1 <= T(0) &&
          std::fabs(r0 * old1 - r1 * old0) <= T(2e-7) * on * rn)
        return {old0, old1};
      T ox = vx * old0 + vy * old1, oy = -vy * old0 + vx * old1;
      T b0 = largest * ox - r0r, b1 = smallest * oy - r1r;
      T x = 0, y = 0;
      T hi = std::sqrt(b0 * b0 + b1 * b1) / radius;
      if (hi > T(0)) {
        T lo_ = mx(mx(hi - largest, std::fabs(b1) / radius - smallest), T(0));
        T alpha = lo_;
        bool done = false;
        for (uint32_t it = 0; it < 8; it++) {
          T i0 = T(1) / (largest + alpha);
          T i1 = smallest + alpha > T(0) ? T(1) / (smallest + alpha) : T(0);
          T tx = ox - (r0r + alpha * ox) * i0, ty = oy - (r1r + alpha * oy) * i1;
          T nrm = std::sqrt(tx * tx + ty * ty);
          if (std::fabs(nrm - radius) <= T(2e-7) * radius || (alpha == T(0) && nrm <= radius)) { x = tx; y = ty; done = true; break; }
          if (nrm > radius) lo_ = alpha; else hi = alpha;
          T slope = tx * tx * i0 + ty * ty * i1;
          T cand = alpha;
          if (slope > T(0)) cand = alpha + (nrm / radius - T(1)) * nrm * nrm / slope;
          alpha = (cand > lo_ && cand < hi) ? cand : T(0.5) * (lo_ + hi);
        }
        if (!done) {
          for (uint32_t it = 0; it < 24; it++) {
            alpha = T(0.5) * (lo_ + hi);
            if (length(V2<T>{b0 / (largest + alpha), b1 / (smallest + alpha)}) > radius) lo_ = alpha; else hi = alpha;
          }
          x = b0 / (largest + hi); y = b1 / (smallest + hi);
        }
      }
      res = {old0 + vx * (x - ox) - vy * (y - oy), old1 + vy * (x - ox) + vx * (y - oy)};
    }
    V2<T> oldv{old0, old1};
    V2<T> relaxed = oldv + omega * (res - oldv);
    T m = length(relaxed);
    return m > radius ? relaxed * (radius / m) : relaxed;
  }
  void solve_unit(uint32_t us, uint32_t row, int a, int b, uint32_t mode, uint32_t it) {  // wgsl.js:3033-3231
    (void)it;
    V3<T> ta[2], tb[2];
    load_tw(a, ta); load_tw(b, tb);
    uint32_t o = row * ROWS;
    V4<T> r0 = rows[o], r1 = rows[o + 1], r2 = rows[o + 2];
    V3<T> ra = xyz(rows[o + 3]), rb = xyz(rows[o + 4]);
    V3<T> n = xyz(r0), t1 = xyz(r1), t2 = xyz(r2);
    V4<T> lam = rows[o + 6];
    T dn = 0, dt1 = 0, dt2 = 0;
    if (mode == 0) {
      dn = lam.x; dt1 = lam.y; dt2 = lam.z;
      for (uint32_t r = 0; r < 3; r++) {
        T d = r == 0 ? dn : (r == 1 ? dt1 : dt2);
        ta[0] += xyz(rows[o + 7 + r * 4]) * d; ta[1] += xyz(rows[o + 8 + r * 4]) * d;
        tb[0] += xyz(rows[o + 9 + r * 4]) * d; tb[1] += xyz(rows[o + 10 + r * 4]) * d;
      }
      V3<T> fra = ra, frb = rb;
      if (EXTRA) { fra = xyz(rows[o + PO]); frb = xyz(rows[o + PO + 1]); }
      V3<T> fn = n * dn, ft = t1 * dt1 + t2 * dt2;
      add_imp(a, -fn - ft, -cross(ra, fn) - cross(fra, ft));
      add_imp(b, fn + ft, cross(rb, fn) + cross(frb, ft));
    } else {
      V3<T> vrel = (tb[0] + cross(tb[1], rb)) - (ta[0] + cross(ta[1], ra));
      T jv = dot(n, vrel);
      T w = vpass ? T(1) : rows[o + 5].w;
      T res = jv + r1.w;
      if (COMPL) res += rows[o + CO].x * lam.x;
      T nl = mx(lam.x + F_OMEGA * (-res * r0.w * w - (T(1) - w) * lam.x), T(0));
      dn = nl - lam.x;
      lam.x = nl;
      ta[0] += xyz(rows[o + 7]) * dn; ta[1] += xyz(rows[o + 8]) * dn;
      tb[0] += xyz(rows[o + 9]) * dn; tb[1] += xyz(rows[o + 10]) * dn;
      add_imp(a, -n * dn, -cross(ra, n * dn));
      add_imp(b, n * dn, cross(rb, n * dn));
      V3<T> fra = ra, frb = rb;
      T lim = r2.w * nl, pooled = nl;
      V2<T> bias{0, 0};
      T role = 1;
      if (EXTRA) {
        role = rows[o + PO].w;
        if (PATCH && role != T(1)) art[A_ULOAD + us] += dn;
        fra = xyz(rows[o + PO]); frb = xyz(rows[o + PO + 1]);
        if (PATCH && role == T(2)) {
          T load = art[A_ULOAD + us];
          lim = r2.w * load;
          pooled = load;
          V3<T> e = xyz(rows[o + PO + 2]);
          if (!vpass) bias = {F_ANCHOR_BETA * dot(t1, e) / dt, F_ANCHOR_BETA * dot(t2, e) / dt};
          rows[o + PO + 2].w = lim;
        }
      }
      if (FMODE) {
        // Coupled (normal, t1, t2) re-solve; overwrites all three impulses (wgsl.js:3081-3097).
        V3<T> vn = (tb[0] + cross(tb[1], rb)) - (ta[0] + cross(ta[1], ra));
        V3<T> vt = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
        V3<T> nlam = contact_block(FMODE, dot(n, vn), V2<T>{dot(t1, vt), dot(t2, vt)}, rows[o + FMO], rows[o + FMO + 1], -r1.w, r2.w,
                                   xyz(lam));
        V3<T> d = nlam - xyz(lam);
        if (PATCH && role != T(1)) art[A_ULOAD + us] += d.x;
        ta[0] += xyz(rows[o + 7]) * d.x + xyz(rows[o + 11]) * d.y + xyz(rows[o + 15]) * d.z;
        ta[1] += xyz(rows[o + 8]) * d.x + xyz(rows[o + 12]) * d.y + xyz(rows[o + 16]) * d.z;
        tb[0] += xyz(rows[o + 9]) * d.x + xyz(rows[o + 13]) * d.y + xyz(rows[o + 17]) * d.z;
        tb[1] += xyz(rows[o + 10]) * d.x + xyz(rows[o + 14]) * d.y + xyz(rows[o + 18]) * d.z;
        V3<T> fn_ = n * d.x, ft = t1 * d.y + t2 * d.z;
        add_imp(a, -fn_ - ft, -cross(ra, fn_) - cross(fra, ft));
        add_imp(b, fn_ + ft, cross(rb, fn_) + cross(frb, ft));
        lam = v4(nlam, lam.w);
      } else if (role != T(0)) {
        V3<T> vr2 = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
        V4<T> ki = rows[o + 5];
        T v1 = dot(t1, vr2) + bias.x, v2 = dot(t2, vr2) + bias.y;
        T kd = ki.x * ki.z - ki.y * ki.y;
        V2<T> lt = friction_pair_solve(ki.z / kd, -ki.y / kd, ki.x / kd, v1, v2, lam.y, lam.z, lim, F_OMEGA);
        dt1 = lt.x - lam.y; dt2 = lt.y - lam.z;
        lam.y = lt.x; lam.z = lt.y;
        ta[0] += xyz(rows[o + 11]) * dt1 + xyz(rows[o + 15]) * dt2;
        ta[1] += xyz(rows[o + 12]) * dt1 + xyz(rows[o + 16]) * dt2;
        tb[0] += xyz(rows[o + 13]) * dt1 + xyz(rows[o + 17]) * dt2;
        tb[1] += xyz(rows[o + 14]) * dt1 + xyz(rows[o + 18]) * dt2;
        V3<T> f = t1 * dt1 + t2 * dt2;
        add_imp(a, -f, -cross(fra, f));
        add_imp(b, f, cross(frb, f));
      }
      const bool art_u = (a >= 0 && body_kind(a) >= 2) || (b >= 0 && body_kind(b) >= 2);
      const bool ell = ANG_ART && ANG && art_u && (rows[o + 19].x > T(0) || rows[o + 19].y > T(0));
      if (ell) {
        // Articulated spin/rolling: after the sliding pair, Steel's block over sliding, spin and rolling (wgsl.js:3113-3141).
        T H5[25];
        uint32_t idx = 0;
        for (uint32_t x = 0; x < 5; x++)
          for (uint32_t y2 = x; y2 < 5; y2++) {
            T hv = (&rows[o + EO + idx / 4].x)[idx % 4];
            H5[x * 5 + y2] = hv; H5[y2 * 5 + x] = hv;
            idx++;
          }
        V3<T> vt = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
        V3<T> wr = tb[1] - ta[1];
        V4<T> la0 = rows[o + 21];
        T g5[5] = {dot(t1, vt), dot(t2, vt), dot(n, wr), dot(t1, wr), dot(t2, wr)};
        T mu5[5] = {r2.w, r2.w, rows[o + 19].x, rows[o + 19].y, rows[o + 19].y};
        T l5[5] = {lam.y, lam.z, la0.x, la0.y, la0.z};
        T nl5[5], d[5];
        ell_block(H5, g5, mu5, l5, pooled, nl5);
        for (uint32_t k = 0; k < 5; k++) d[k] = nl5[k] - l5[k];
        ta[0] += xyz(rows[o + 11]) * d[0] + xyz(rows[o + 15]) * d[1] + xyz(rows[o + 22]) * d[2] + xyz(rows[o + 26]) * d[3] + xyz(rows[o + 30]) * d[4];
        ta[1] += xyz(rows[o + 12]) * d[0] + xyz(rows[o + 16]) * d[1] + xyz(rows[o + 23]) * d[2] + xyz(rows[o + 27]) * d[3] + xyz(rows[o + 31]) * d[4];
        tb[0] += xyz(rows[o + 13]) * d[0] + xyz(rows[o + 17]) * d[1] + xyz(rows[o + 24]) * d[2] + xyz(rows[o + 28]) * d[3] + xyz(rows[o + 32]) * d[4];
        tb[1] += xyz(rows[o + 14]) * d[0] + xyz(rows[o + 18]) * d[1] + xyz(rows[o + 25]) * d[2] + xyz(rows[o + 29]) * d[3] + xyz(rows[o + 33]) * d[4];
        V3<T> f = t1 * d[0] + t2 * d[1];
        V3<T> tq = n * d[2] + t1 * d[3] + t2 * d[4];
        add_imp(a, -f, -cross(fra, f) - tq);
        add_imp(b, f, cross(frb, f) + tq);
        lam.y = nl5[0]; lam.z = nl5[1];
        rows[o + 21] = {nl5[2], nl5[3], nl5[4], T(0)};
      }
      rows[o + 6] = lam;
      if (ANG) {
        // Spin within mu_t N, rolling within a disk of mu_r N; under friction patches only the region's first anchor
        // carries them, with the pooled load (wgsl.js:3144-3173).
        V4<T> g = rows[o + 19];
        bool aon = !ell && (!PATCH || role == T(1) || (role == T(2) && rows[o + PO + 1].w == T(0)));
        if (aon && (g.x > T(0) || g.y > T(0))) {
          V4<T> la = rows[o + 21];
          V3<T> wrel = tb[1] - ta[1];
          T ns = clampv(la.x - dot(n, wrel) * g.z, -g.x * pooled, g.x * pooled);
          T ds = ns - la.x;
          ta[0] += xyz(rows[o + 22]) * ds; ta[1] += xyz(rows[o + 23]) * ds;
          tb[0] += xyz(rows[o + 24]) * ds; tb[1] += xyz(rows[o + 25]) * ds;
          V3<T> w2 = tb[1] - ta[1];
          V4<T> ri = rows[o + 20];
          T u1 = dot(t1, w2), u2 = dot(t2, w2);
          V2<T> lr = V2<T>{la.y, la.z} - V2<T>{ri.x * u1 + ri.y * u2, ri.y * u1 + ri.z * u2};
          T rl = g.y * pooled;
          if (length(lr) > rl) lr = lr * (rl / mx(length(lr), T(1e-12)));
          T d1 = lr.x - la.y, d2 = lr.y - la.z;
          ta[0] += xyz(rows[o + 26]) * d1 + xyz(rows[o + 30]) * d2; ta[1] += xyz(rows[o + 27]) * d1 + xyz(rows[o + 31]) * d2;
          tb[0] += xyz(rows[o + 28]) * d1 + xyz(rows[o + 32]) * d2; tb[1] += xyz(rows[o + 29]) * d1 + xyz(rows[o + 33]) * d2;
          rows[o + 21] = {ns, lr.x, lr.y, T(0)};
          V3<T> tq = n * ds + t1 * d1 + t2 * d2;
          add_imp(a, V3<T>{0, 0, 0}, -tq);
          add_imp(b, V3<T>{0, 0, 0}, tq);
        }
      }
      if (TORS && row == S_UTP[us] && rows[S_USTART[us] * ROWS + TO].w > T(0)) unit_torsion(us, a, b, ta, tb);
    }
    if (a >= 0) { tw[(uint32_t)a * 2] = v4(ta[0], T(0)); tw[(uint32_t)a * 2 + 1] = v4(ta[1], T(0)); }
    if (b >= 0) { tw[(uint32_t)b * 2] = v4(tb[0], T(0)); tw[(uint32_t)b * 2 + 1] = v4(tb[1], T(0)); }
  }
  // Contact torsion after the unit's last colored row (wgsl.js:3185-3228): sliding and spin share the members'
  // Coulomb budget mu * N; when sliding exceeds it (later normals lowered the load) the members' tangent pairs are
  // scaled onto it, then the spin row is relaxed within r * (budget - sliding).
  void unit_torsion(uint32_t us, int a, int b, V3<T> ta[2], V3<T> tb[2]) {
    uint32_t s0 = S_USTART[us], s1 = std::min(s0 + S_USIZE[us], MAXC), L = s0 * ROWS;
    T mu = rows[L + TO].w;
    T load = 0, slide = 0;
    for (uint32_t m = s0; m < s1; m++) {
      uint32_t om = m * ROWS;
      if (rows[om + TO + 4].w < T(0.5)) continue;
      V4<T> lm = rows[om + 6];
      load += mx(lm.x, T(0));
      if (!EXTRA || rows[om + PO].w != T(0)) slide += length(V2<T>{lm.y, lm.z});
    }
    T budget = mu * load;
    if (slide > budget && slide > T(0)) {
      T scale = budget / slide;
      for (uint32_t m = s0; m < s1; m++) {
        uint32_t om = m * ROWS;
        if (rows[om + TO + 4].w < T(0.5) || (EXTRA && rows[om + PO].w == T(0))) continue;
        V4<T> lm = rows[om + 6];
        T d1 = lm.y * (scale - T(1)), d2 = lm.z * (scale - T(1));
        ta[0] += xyz(rows[om + 11]) * d1 + xyz(rows[om + 15]) * d2; ta[1] += xyz(rows[om + 12]) * d1 + xyz(rows[om + 16]) * d2;
        tb[0] += xyz(rows[om + 13]) * d1 + xyz(rows[om + 17]) * d2; tb[1] += xyz(rows[om + 14]) * d1 + xyz(rows[om + 18]) * d2;
        V3<T> fa = xyz(rows[om + 3]), fb = xyz(rows[om + 4]);
        if (EXTRA) { fa = xyz(rows[om + PO]); fb = xyz(rows[om + PO + 1]); }
        V3<T> f = xyz(rows[om + 1]) * d1 + xyz(rows[om + 2]) * d2;
        add_imp(a, -f, -cross(fa, f));
        add_imp(b, f, cross(fb, f));
        rows[om + 6] = {lm.x, lm.y * scale, lm.z * scale, lm.w};
      }
      slide = budget;
    }
    T bound = rows[L + TO + 3].w * mx(budget - slide, T(0));
    V3<T> ng = xyz(rows[L + TO + 4]);
    T ls = rows[L + TO + 2].w;
    T nls = clampv(ls - F_OMEGA * dot(ng, tb[1] - ta[1]) * rows[L + TO + 1].w, -bound, bound);
    T ds = nls - ls;
    if (ds != T(0)) {
      ta[0] += xyz(rows[L + TO]) * ds; ta[1] += xyz(rows[L + TO + 1]) * ds;
      tb[0] += xyz(rows[L + TO + 2]) * ds; tb[1] += xyz(rows[L + TO + 3]) * ds;
      rows[L + TO + 2].w = nls;
      add_imp(a, V3<T>{0, 0, 0}, -ng * ds);
      add_imp(b, V3<T>{0, 0, 0}, ng * ds);
    }
  }
  void set_link_twist(uint32_t l, SV<T> v) {
    uint32_t b = (uint32_t)link_body(l);
    tw[b * 2] = v4(v.l, T(0)); tw[b * 2 + 1] = v4(v.a, T(0));
  }
  void art_twists(uint32_t a) {
    uint32_t ai[4];
    art_info(a, ai);
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k;
      int par = link_parent(l);
      SV<T> v = sv0<T>();
      if (par >= 0) v = mshift(ldsv(lo((uint32_t)par) + L_V), link_r(l));
      for (uint32_t r = 0; r < link_ndof(l); r++) v = sadd(v, sscale(ldS(l, r), aqd[link_dof(l) + r]));
      stsv(lo(l) + L_V, v);
      set_link_twist(l, v);
    }
  }
  void art_apply(uint32_t a) {  // wgsl.js:3254-3299
    uint32_t ai[4];
    art_info(a, ai);
    for (uint32_t kk = 0; kk < ai[1]; kk++) {
      uint32_t l = ai[0] + ai[1] - 1 - kk, L = lo(l);
      uint32_t n = link_ndof(l);
      SV<T> p = ldsv(L + L_P);
      T u[6];
      for (uint32_t r = 0; r < n; r++) { u[r] = art[L + L_TAU + r] - sdot(ldS(l, r), p); art[L + L_UU + r] = u[r]; }
      int par = link_parent(l);
      if (par >= 0) {
        SV<T> pp = p;
        for (uint32_t r = 0; r < n; r++) {
          T du = 0;
          for (uint32_t c = 0; c < n; c++) du += ldDI(l, r, c) * u[c];
          pp = sadd(pp, sscale(ldU(l, r), du));
        }
        uint32_t P0 = lo((uint32_t)par) + L_P;
        stsv(P0, sadd(ldsv(P0), fshift(pp, link_r(l))));
      }
    }
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k, L = lo(l);
      int par = link_parent(l);
      SV<T> dv = sv0<T>();
      if (par >= 0) dv = mshift(ldsv(lo((uint32_t)par) + L_DV), link_r(l));
      uint32_t n = link_ndof(l);
      T dq[6];
      for (uint32_t r = 0; r < n; r++) {
        dq[r] = 0;
        for (uint32_t c = 0; c < n; c++) dq[r] += ldDI(l, r, c) * (art[L + L_UU + c] - sdot(ldU(l, c), dv));
      }
      for (uint32_t r = 0; r < n; r++) { dv = sadd(dv, sscale(ldS(l, r), dq[r])); aqd[link_dof(l) + r] += dq[r]; }
      stsv(L + L_DV, dv);
      SV<T> v = sadd(ldsv(L + L_V), dv);
      stsv(L + L_V, v);
      set_link_twist(l, v);
      stsv(L + L_P, sv0<T>());
      for (uint32_t r = 0; r < JM; r++) art[L + L_TAU + r] = 0;
    }
  }
  // subgroupAdd of the reference's 32-lane subgroup path (wgsl.js:3305); WGSL leaves its order open, this
  // build fixes it as an xor butterfly over strides 16, 8, 4, 2, 1 (lane i adds lane i^s's value).
  static T lane_sum32(T (&v)[32]) {
    for (uint32_t s = 16; s >= 1; s >>= 1) {
      T nv[32];
      for (uint32_t i = 0; i < 32; i++) nv[i] = v[i] + v[i ^ s];
      for (uint32_t i = 0; i < 32; i++) v[i] = nv[i];
    }
    return v[0];
  }
  // Dense same-articulation rows, subgroup form (wgsl.js:3306-3364, SG = 1): lane l owns dofs l, l+32, ...
  // and sums its own products in dof order before lane_sum32.
  bool art_dense(uint32_t a, uint32_t mode, uint32_t it) {
    (void)it;
    uint32_t nd = std::min(S_META[3], MAXDENSE);
    uint32_t ai[4];
    art_info(a, ai);
    const uint32_t n = ai[3];
    bool touched = false;
    for (uint32_t i = 0; i < nd; i++) {
      uint32_t o = A_DENSE + i * DPS;
      if (tu(art[o]) != a) continue;
      touched = true;
      uint32_t jo = o + DHDR, yo = o + DHDR + 3 * NDA;
      V3<T> lam{art[o + 9], art[o + 10], art[o + 11]};
      if (art[o + 21] > T(0.5)) {
        // Bilateral rows (mimics): unclamped, one at a time; no warm impulse (wgsl.js:3319-3332).
        if (mode == 0) continue;
        T nlam[3] = {lam.x, lam.y, lam.z};
        for (uint32_t r = 0; r < tu(art[o + 21]); r++) {
          T part[32];
          for (uint32_t l = 0; l < 32; l++) {
            part[l] = T(0);
            for (uint32_t k = l; k < n; k += 32) part[l] += art[jo + r * NDA + k] * aqd[ai[2] + k];
          }
          T jv = lane_sum32(part);
          T rhs = vpass ? T(0) : (r == 0 ? art[o + 4] : (r == 1 ? art[o + 22] : art[o + 23]));
          T dl = -F_OMEGA * (jv + rhs) * art[o + 5 + r];
          nlam[r] += dl;
          for (uint32_t k = 0; k < n; k++) aqd[ai[2] + k] += art[yo + r * NDA + k] * dl;
        }
        art[o + 9] = nlam[0]; art[o + 10] = nlam[1]; art[o + 11] = nlam[2];
        continue;
      }
      if (mode == 0) {
        for (uint32_t k = 0; k < n; k++) aqd[ai[2] + k] += art[yo + k] * lam.x + art[yo + NDA + k] * lam.y + art[yo + 2 * NDA + k] * lam.z;
        continue;
      }
      T part[32];
      for (uint32_t l = 0; l < 32; l++) {
        part[l] = T(0);
        for (uint32_t k = l; k < n; k += 32) part[l] += art[jo + k] * aqd[ai[2] + k];
      }
      T jv = lane_sum32(part);
      T w = vpass ? T(1) : art[o + 24];
      T nl = mx(lam.x + F_OMEGA * (-(jv + art[o + 4]) * art[o + 5] * w - (T(1) - w) * lam.x), T(0));
      T dn = nl - lam.x;
      T p1[32], p2[32];
      for (uint32_t l = 0; l < 32; l++) {
        p1[l] = T(0); p2[l] = T(0);
        for (uint32_t k = l; k < n; k += 32) {
          aqd[ai[2] + k] += art[yo + k] * dn;
          p1[l] += art[jo + NDA + k] * aqd[ai[2] + k];
          p2[l] += art[jo + 2 * NDA + k] * aqd[ai[2] + k];
        }
      }
      T v1 = lane_sum32(p1), v2 = lane_sum32(p2);
      T lim = art[o + 3] * nl;
      T kd = art[o + 6] * art[o + 8] - art[o + 7] * art[o + 7];
      V2<T> lt = friction_pair_solve(art[o + 8] / kd, -art[o + 7] / kd, art[o + 6] / kd, v1, v2, lam.y, lam.z, lim, F_OMEGA);
      T d1 = lt.x - lam.y, d2 = lt.y - lam.z;
      for (uint32_t k = 0; k < n; k++) aqd[ai[2] + k] += art[yo + NDA + k] * d1 + art[yo + 2 * NDA + k] * d2;
      art[o + 9] = nl; art[o + 10] = lt.x; art[o + 11] = lt.y;
    }
    return touched;
  }
  void dense_phase(uint32_t mode, uint32_t it) {
    for (uint32_t a = 0; a < NA; a++) if (art_dense(a, mode, it)) art_twists(a);
  }
  void art_contact_impulses(uint32_t a) {
    uint32_t ai[4];
    art_info(a, ai);
    bool hit = false;
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k, b = (uint32_t)link_body(l);
      V3<T> f = xyz(imp[b * 2]), t = xyz(imp[b * 2 + 1]);
      if (vne0(f) || vne0(t)) {
        hit = true;
        stsv(lo(l) + L_P, SV<T>{-t, -f});
        imp[b * 2] = {0, 0, 0, 0}; imp[b * 2 + 1] = {0, 0, 0, 0};
      }
    }
    if (hit) art_apply(a);
  }
  void art_dof_rows(uint32_t a, uint32_t l, uint32_t r, uint32_t d) {  // wgsl.js:3421-3468, drive, friction, limit rows
    T hinv = mx(art[A_HINV + d], T(1e-12));
    V4<T> dr = dof_drive(d);
    if (DRIVE_ROWS && driven_dof(d) && !(vpass && !VEL_DRIVE_ACTIVE)) {
      // drive_mode rows: a PhysX-style drive row, cold-started each step, not relaxed (wgsl.js:3425-3438).
      T q = art[A_Q + link_qs(l) + r];
      T pd = vpass ? dt * art[A_SNAP + NB * 8 + d] : T(0);
      T ca = dt * (dt * dr.y + dr.z);
      T x = T(1) / (T(1) + ca * hinv);
      T bias = x * dt * dr.z * art[A_TGTV + d] + dr.y * x * dt * (art[A_TGT + d] - q - pd);
      T cap = dr.w > T(0) ? dr.w * dt : T(1e20);
      T lam = art[A_DRV + d];
      T nl = mn(mx(lam * (T(1) - x) - x * ca * aqd[d] + bias, -cap), cap);
      if (nl != lam) {
        art[A_DRV + d] = nl;
        art[lo(l) + L_TAU + r] = nl - lam;
        art_apply(a);
      }
    }
    // Joint friction: a velocity row with |impulse| <= friction * dt (plus the load term).
    T fr = dof_passive(d).y;
    if (JFL) fr += art[A_JFL + d];
    if (fr > T(0)) {
      T lam = art[A_FRIC + d];
      T nl = mn(mx(lam - F_OMEGA * aqd[d] / hinv, -fr * dt), fr * dt);
      if (nl != lam) {
        art[A_FRIC + d] = nl;
        art[lo(l) + L_TAU + r] = nl - lam;
        art_apply(a);
      }
    }
    V4<T> lim = dof_limit(d);
    if (art[A_LG + d] > T(0.5)) {
      T q = art[A_Q + link_qs(l) + r];
      for (uint32_t side = 0; side < 2; side++) {
        T sg = side == 0 ? T(1) : T(-1);
        T gap = side == 0 ? q - lim.x : lim.y - q;
        T lam = art[A_LIM + 2 * d + side];
        T rhs = gap / dt;
        if (gap < T(0)) rhs = vpass ? T(0) : mx(beta * gap / dt, -max_depen);
        T nl = mx(lam - F_OMEGA * (sg * aqd[d] + rhs) / hinv, T(0));
        if (nl == lam) continue;
        art[A_LIM + 2 * d + side] = nl;
        art[lo(l) + L_TAU + r] = sg * (nl - lam);
        art_apply(a);
      }
    }
  }
  void art_propagate_rows(uint32_t a) {
    uint32_t ai[4];
    art_info(a, ai);
    uint32_t n = tu(art[A_RLN + a]);
    for (uint32_t i = 0; i < n; i++) {
      uint32_t d = tu(art[A_RL + ai[2] + i]);
      uint32_t l = tu(dof_limit(d).w);
      art_dof_rows(a, l, d - link_dof(l), d);
    }
  }
  // Velocity limits, last in a sweep: a stateless push of each active scalar dof back onto |qd| <= limit
  // through the articulation (wgsl.js:3413-3419, 3462-3467).
  void art_velocity_limits(uint32_t a) {
    if (!(VLIM_ON && art[A_AVL + a] > T(0.5))) return;
    uint32_t ai[4];
    art_info(a, ai);
    for (uint32_t k = 0; k < ai[1]; k++) {
      uint32_t l = ai[0] + k, t = link_type(l);
      if (t != JT_REVOLUTE && t != JT_PRISMATIC && t != JT_D6) continue;
      for (uint32_t r = 0; r < link_ndof(l); r++) {
        uint32_t d = link_dof(l) + r;
        T hinv = mx(art[A_HINV + d], T(1e-12));
        V4<T> pv = dof_passive(d);
        if (pv.z >= T(0) && art[A_VLG + d] > T(0.5)) {
          T over = aqd[d] - mn(mx(aqd[d], -pv.z), pv.z);
          if (over != T(0)) {
            art[lo(l) + L_TAU + r] = -over / hinv;
            art_apply(a);
          }
        }
      }
    }
  }
  void sweep(uint32_t mode, uint32_t ncol, uint32_t it, const std::vector<uint32_t>& wcs, const std::vector<uint32_t>& slot, uint32_t nu) {
    for (uint32_t a = 0; a < NA; a++) art_propagate_rows(a);
    if (PATCH && mode == 1) {
      for (uint32_t us = 0; us < nu; us++) {
        uint32_t s0 = S_USTART[us];
        T load = 0;
        for (uint32_t m = s0; m < s0 + S_USIZE[us]; m++) if (rows[m * ROWS + PO].w != T(1)) load += rows[m * ROWS + 6].x;
        art[A_ULOAD + us] = load;
      }
    }
    dense_phase(1, it);
    for (uint32_t c = 0; c < NCOL && c < ncol; c++)
      for (uint32_t u = wcs[c]; u < wcs[c + 1]; u++) solve_unit(slot[u * 4 + 1], slot[u * 4], (int)slot[u * 4 + 2], (int)slot[u * 4 + 3], mode, it);
    for (uint32_t u = wcs[NCOL]; u < wcs[NCOL + 1]; u++) {
      uint32_t uu = std::min(u, MAXC - 1);
      solve_unit(slot[uu * 4 + 1], slot[uu * 4], (int)slot[uu * 4 + 2], (int)slot[uu * 4 + 3], mode, it);
    }
    for (uint32_t a = 0; a < NA; a++) {
      art_contact_impulses(a);
      art_velocity_limits(a);
    }
    if (VLIM_FREE) for (uint32_t b = 0; b < NB; b++) body_velocity_limits(b);
  }
  // Free-body velocity limits: a per-axis box on v (clamped) and w (pushed back through the body's inverse
  // inertia), stateless (wgsl.js:3586-3610).
  void body_velocity_limits(uint32_t b) {
    if (body_kind((int)b) != 1 || (body_flags((int)b) & 1) != 0) return;
    V4<T> lm = body_limits((int)b);
    uint32_t g = A_VLG + ND + b * 6;
    if (lm.y > T(0)) {
      V3<T> v = xyz(tw[b * 2]);
      for (uint32_t k = 0; k < 3; k++) if (art[g + k] > T(0.5)) v[k] = mn(mx(v[k], -lm.y), lm.y);
      tw[b * 2] = v4(v, T(0));
    }
    if (lm.z > T(0)) {
      V3<T> w = xyz(tw[b * 2 + 1]);
      for (uint32_t k = 0; k < 3; k++) {
        T over = w[k] - mn(mx(w[k], -lm.z), lm.z);
        if (over == T(0) || art[g + 3 + k] < T(0.5)) continue;
        V3<T> e{0, 0, 0};
        e[k] = T(1);
        M3<T> R = qmat(poses[P_BODY + b * 2 + 1]);
        V3<T> iw = mul(R, mul(body_inv_inertia((int)b), mul(transpose(R), e)));
        w = w - iw * (over / iw[k]);
      }
      tw[b * 2 + 1] = v4(w, T(0));
    }
  }
  V3<T> gyroscopic_velocity(const M3<T>& inertia, V3<T> u0, T h0) const {  // wgsl.js:2543-2562
    auto inverse3 = [](const M3<T>& m) {
      V3<T> a = m.c[0], b = m.c[1], c = m.c[2];
      V3<T> r0 = cross(b, c), r1 = cross(c, a), r2 = cross(a, b);
      T det = dot(a, r0);
      M3<T> t; t.c[0] = r0; t.c[1] = r1; t.c[2] = r2;
      return mscale(transpose(t), T(1) / det);
    };
    M3<T> a = mscale(inertia, T(1) / mx(inertia.c[0].x, mx(inertia.c[1].y, inertia.c[2].z)));
    M3<T> inv = inverse3(a);
    T speed = std::sqrt(mx(dot(u0, mul(a, u0)) * (inv.c[0].x + inv.c[1].y + inv.c[2].z), T(0)));
    uint32_t n = (uint32_t)clampv(std::ceil(T(2) * std::fabs(h0) * speed), T(1), T(32));
    T h = h0 / (T)n;
    V3<T> w = u0;
    for (uint32_t m = 0; m < n; m++) {
      V3<T> u = w;
      T energy = dot(u, mul(a, u));
      for (uint32_t it = 0; it < 8; it++) {
        M3<T> S = mscale(T(0.5) * h, skew3(mul(a, T(0.5) * (u + w))));
        V3<T> c = mul(inverse3(msub(a, S)), mul(madd(a, S), u));
        V3<T> d = c - w;
        w = c;
        if (dot(d, mul(a, d)) <= T(1e-12) * energy) break;
      }
    }
    return w;
  }
  V3<T> body_motion(int b, V3<T> x) const {  // wgsl.js:3472-3489
    if (b < 0) return {0, 0, 0};
    V3<T> p0 = xyz(poses[P_PREV + (uint32_t)b * 2]);
    V4<T> q0 = poses[P_PREV + (uint32_t)b * 2 + 1];
    V3<T> p1 = xyz(poses[P_BODY + (uint32_t)b * 2]);
    V4<T> q1 = poses[P_BODY + (uint32_t)b * 2 + 1];
    V4<T> r = qmul(q1, qconj(q0));
    if (r.w < T(0)) r = {-r.x, -r.y, -r.z, -r.w};
    T sn = length(xyz(r));
    V3<T> ang = T(2) * xyz(r);
    V3<T> tr = p1 - p0;
    V3<T> lin = tr;
    if (sn > T(1e-6)) {
      T half = tatan2<T>(sn, r.w);
      V3<T> axis = xyz(r) / sn;
      ang = (T(2) * half) * axis;
      T sc = half * r.w / sn;
      lin = sc * tr + ((T(1) - sc) * dot(axis, tr)) * axis;
    }
    return lin + cross(ang, x - T(0.5) * (p0 + p1));
  }
  void solve() {
    uint32_t nu = S_META[0], ncol = std::min(S_META[1], NCOL);
    std::vector<uint32_t> wcs(NCOL + 2);
    for (uint32_t i = 0; i < NCOL + 2; i++) wcs[i] = S_CSTART[i];
    for (uint32_t i = 0; i < ND; i++) aqd[i] = art[A_QDH + i];
    if (SLEEP)
      for (uint32_t a = 0; a < NA; a++) {
        if (!node_asleep((int)(NB + a))) continue;
        uint32_t ai[4];
        art_info(a, ai);
        for (uint32_t d = ai[2]; d < ai[2] + ai[3]; d++) aqd[d] = T(0);
      }
    for (uint32_t b = 0; b < NB; b++) {
      uint32_t kind = body_kind((int)b);
      imp[b * 2] = {0, 0, 0, 0}; imp[b * 2 + 1] = {0, 0, 0, 0};
      if (kind == 1) {
        uint32_t fl = body_flags((int)b);
        V3<T> v = xyz(vel[b * 2]) + ((node_asleep((int)b) || fl != 0) ? V3<T>{0, 0, 0} : g * dt);
        V3<T> w = xyz(vel[b * 2 + 1]);
        uint32_t bfo = A_BF + b * 8;
        if ((fl & 1) == 0 && (vne0(ld3(bfo)) || vne0(ld3(bfo + 4)))) {
          M3<T> R = qmat(poses[P_BODY + b * 2 + 1]);
          v += ld3(bfo) * body_inv_mass((int)b) * dt;
          w += mul(R, mul(body_inv_inertia((int)b), mul(transpose(R), ld3(bfo + 4)))) * dt;
        }
        M3<T> I = body_inertia((int)b);
        bool iso = I.c[0].x == I.c[1].y && I.c[1].y == I.c[2].z && I.c[0].y == T(0) && I.c[0].z == T(0) && I.c[1].z == T(0);
        if ((fl & 1) == 0 && !iso && vne0(w)) {
          M3<T> R = qmat(poses[P_BODY + b * 2 + 1]);
          w = mul(R, gyroscopic_velocity(I, mul(transpose(R), w), dt));
        }
        tw[b * 2] = v4(v, T(0)); tw[b * 2 + 1] = v4(w, T(0));
      }
    }
    for (uint32_t i = 0; i < 2 * ND; i++) art[A_LIM + i] = 0;
    for (uint32_t i = 0; i < ND; i++) { art[A_FRIC + i] = 0; art[A_DRV + i] = 0; }
    uint32_t nslot = std::min(S_CSTART[NCOL + 1], MAXC);
    std::vector<uint32_t> slot(MAXC * 4, 0);
    for (uint32_t u = 0; u < nslot; u++) {
      uint32_t row = S_PORD[u], us = S_RUNIT[row];
      uint64_t key = S_UKEY[us];
      slot[u * 4] = row; slot[u * 4 + 1] = us; slot[u * 4 + 2] = (uint32_t)key_a(key); slot[u * 4 + 3] = (uint32_t)key_b(key);
    }
    for (uint32_t a = 0; a < NA; a++) art_twists(a);
    dense_phase(0, 0);
    sweep(0, ncol, 0, wcs, slot, nu);
    for (uint32_t it = 0; it < ITERS; it++) {
      if (NEST) momentum_step(ncol, it, wcs, slot);
      sweep(1, ncol, it, wcs, slot, nu);
    }
    if (VELITERS) velocity_pass(ncol, wcs, slot, nu);
    uint32_t nc = S_USTART[nu];
    if (WH) warm_store(std::min(nc, MAXC));
    if (PATCH) for (uint32_t b = 0; b < NB * 2; b++) poses[P_PREV + b] = poses[P_BODY + b];
    for (uint32_t b = 0; b < NB; b++) {
      if (body_kind((int)b) != 1) continue;
      if (node_asleep((int)b)) { vel[b * 2] = {0, 0, 0, 0}; vel[b * 2 + 1] = {0, 0, 0, 0}; continue; }
      V3<T> v = xyz(tw[b * 2]), w = xyz(tw[b * 2 + 1]);
      T damp = (body_flags((int)b) & 1) != 0 ? T(1) : T(1) - body_limits((int)b).x * dt;
      vel[b * 2] = v4(v, T(0)); vel[b * 2 + 1] = v4(w * damp, T(0));
      if (VELITERS) { v = ld3(A_SNAP + b * 8); w = ld3(A_SNAP + b * 8 + 4); }
      V3<T> c = xyz(poses[P_BODY + b * 2]) + v * dt;
      V4<T> q = poses[P_BODY + b * 2 + 1];
      q = normalize4(q + (T(0.5) * dt) * qmul(v4(w, T(0)), q));
      poses[P_BODY + b * 2] = v4(c, T(0));
      poses[P_BODY + b * 2 + 1] = q;
    }
    for (uint32_t i = 0; i < ND; i++) {
      art[A_QD + i] = aqd[i];
      if (VELITERS) aqd[i] = art[A_SNAP + NB * 8 + i];
    }
    for (uint32_t l = 0; l < NL; l++) {
      if (link_type(l) != JT_FREE || link_parent(l) >= 0) continue;
      T k = T(1) - body_limits(link_body(l)).x * dt;
      uint32_t d = link_dof(l);
      art[A_QD + d + 3] *= k; art[A_QD + d + 4] *= k; art[A_QD + d + 5] *= k;
    }
    for (uint32_t a = 0; a < NA; a++) {
      uint32_t ai[4];
      art_info(a, ai);
      for (uint32_t k = 0; k < ai[1]; k++) {
        uint32_t l = ai[0] + k, qs = link_qs(l), d = link_dof(l), t = link_type(l);
        if (t == JT_REVOLUTE || t == JT_PRISMATIC || t == JT_D6)
          for (uint32_t r = 0; r < link_ndof(l); r++) art[A_Q + qs + r] += aqd[d + r] * dt;
        if (t == JT_BALL) {
          V3<T> w{aqd[d], aqd[d + 1], aqd[d + 2]};
          V4<T> q{art[A_Q + qs], art[A_Q + qs + 1], art[A_Q + qs + 2], art[A_Q + qs + 3]};
          q = normalize4(q + (T(0.5) * dt) * qmul(v4(w, T(0)), q));
          art[A_Q + qs] = q.x; art[A_Q + qs + 1] = q.y; art[A_Q + qs + 2] = q.z; art[A_Q + qs + 3] = q.w;
        }
        if (t == JT_FREE) {
          for (uint32_t c = 0; c < 3; c++) art[A_Q + qs + c] += aqd[d + c] * dt;
          V3<T> w{aqd[d + 3], aqd[d + 4], aqd[d + 5]};
          V4<T> q{art[A_Q + qs + 3], art[A_Q + qs + 4], art[A_Q + qs + 5], art[A_Q + qs + 6]};
          q = normalize4(q + (T(0.5) * dt) * qmul(v4(w, T(0)), q));
          art[A_Q + qs + 3] = q.x; art[A_Q + qs + 4] = q.y; art[A_Q + qs + 5] = q.z; art[A_Q + qs + 6] = q.w;
        }
      }
      art_fk(a);
      for (uint32_t k = 0; k < ai[1]; k++) {
        uint32_t l = ai[0] + k, b = (uint32_t)link_body(l);
        poses[P_BODY + b * 2] = v4(ld3(lo(l) + L_COM), T(0));
        poses[P_BODY + b * 2 + 1] = link_rot(l);
      }
    }
    if (PATCH) {
      uint32_t ab = (1 - parity) * AH;
      for (uint32_t i = 0; i < nc; i++) {
        uint32_t o = i * ROWS;
        if (rows[o + PO].w != T(2)) continue;
        uint32_t pair = get_u32(rows[o + PO + 3].w);  // anchor-table pair id; the bodies from the row's unit
        uint64_t ukey = S_UKEY[S_RUNIT[i]];
        int a = key_a(ukey), b = key_b(ukey);
        V3<T> fra = xyz(rows[o + PO]), frb = xyz(rows[o + PO + 1]);
        uint32_t k = tu(rows[o + PO + 1].w);
        V4<T> ex = rows[o + PO + 2];
        T lim = ex.w;
        if (lim <= T(0)) continue;
        V3<T> n = xyz(rows[o]), t1 = xyz(rows[o + 1]), t2 = xyz(rows[o + 2]);
        V3<T> ta[2], tb[2];
        if (VELITERS) { snap_tw(a, ta); snap_tw(b, tb); } else { load_tw(a, ta); load_tw(b, tb); }
        V3<T> vrel = (tb[0] + cross(tb[1], frb)) - (ta[0] + cross(ta[1], fra));
        V4<T> lam = rows[o + 6];
        V3<T> lt = t1 * lam.y + t2 * lam.z;
        T r = mx(mn(body_radius(a), body_radius(b)), T(1e-4));
        if (length(lt) >= (T(1) - T(1e-5)) * lim && dot(vrel, lt) < -lim * T(1e-5) * r / dt) continue;
        V4<T> qa{0, 0, 0, 1};
        V3<T> ca{0, 0, 0};
        if (a >= 0) { qa = poses[P_BODY + (uint32_t)a * 2 + 1]; ca = xyz(poses[P_BODY + (uint32_t)a * 2]); }
        V3<T> ll = xyz(rows[o + PO + 3]);
        V3<T> x = ca + qrot(qa, ll);
        V3<T> e = xyz(ex) + body_motion(b, x) - body_motion(a, x);
        e -= n * dot(n, e);
        V3<T> el = qinvrot(qa, e);
        uint32_t hk = anchor_hash(pair, k);
        uint32_t h = hk % AH;
        for (uint32_t probe = 0; probe < 32; probe++) {
          if (akey[ab + h] == 0) {
            akey[ab + h] = hk;
            AnchorVal<T>& av = aval[ab + h];
            av.nl = qinvrot(qa, n); av.il = qinvrot(qa, lt);
            av.pair = pair; av.k = k; av.ll = ll; av.el = el;
            break;
          }
          h = (h + 1) % AH;
        }
      }
    }
    stats[1] = std::max(stats[1], counters[1]);
    stats[2] = std::max(stats[2], ncol);
    stats[3] = std::max(stats[3], nu);
    stats[4] = std::max(stats[4], counters[0]);
    stats[5] = counters[1];
    stats[6] = ncol;
    stats[8] = S_META[5];
    stats[7] = std::max(stats[7], S_META[3]);
    counters[0] = 0; counters[1] = 0;
  }

  // Reduction inputs and survivors of the last prepare, per sorted contact (the OracleOut red/redf export).
  // The survivors are the contacts S_KEPT lists; the listing is S_CORDER over the units before compaction.
  void export_reduction(uint32_t* red, float* redf) const {
    uint32_t nc = std::min(counters[1], MAXC);
    for (uint32_t c = 0; c < nc; c++) {
      for (int k = 0; k < 8; k++) red[c * 8 + k] = 0;
      red[c * 8] = 0xffffffffu; red[c * 8 + 7] = c;
      for (int k = 0; k < 4; k++) redf[c * 4 + k] = 0.f;
    }
    uint32_t nu = S_META[0];
    uint32_t nl = 0;
    for (uint32_t u = 0; u < nu; u++) nl += S_USIZE_TMP[u];
    nl = std::min(nl, MAXC);
    std::vector<uint32_t> unit_start(nu + 1, 0);
    for (uint32_t u = 0; u < nu; u++) unit_start[u + 1] = unit_start[u] + S_USIZE_TMP[u];
    for (uint32_t p = 0, u = 0; p < nl; p++) {
      while (u + 1 < nu && unit_start[u + 1] <= p) u++;
      uint32_t c = S_CORDER[p];
      BPC g = bp_contact(c);
      red[c * 8] = u; red[c * 8 + 1] = p; red[c * 8 + 2] = g.k[0]; red[c * 8 + 3] = g.k[1]; red[c * 8 + 4] = g.k[2];
      red[c * 8 + 5] = g.pk;
      redf[c * 4] = (float)g.pos.x; redf[c * 4 + 1] = (float)g.pos.y; redf[c * 4 + 2] = (float)g.gap;
    }
    for (uint32_t u = 0; u < nu; u++)  // unit u's survivors sit at its listing start (S_OSTART) in S_KEPT
      for (uint32_t j = 0; j < S_USIZE[u] && S_OSTART[u] + j < MAXC; j++) red[S_KEPT[S_OSTART[u] + j] * 8 + 6] = 1;
  }
  void alloc_scratch() {
    pairs.reserve(2 * MAXP);
    craw.reserve(MAXC);
    rows.assign((size_t)MAXC * ROWS + 2 * MAXC, V4<T>{0, 0, 0, 0});
    auto z = [](auto& v, size_t n) { v.assign(n, 0u); };
    z(S_HKEY, UH); z(S_HCNT, UH); z(S_SLOT2U, UH); z(S_CSLOT, MAXC); z(S_CLOCAL, MAXC); z(S_USIZE_TMP, MAXU); z(S_UKEY, MAXU);
    z(S_USIZE, MAXU); z(S_USTART, MAXU + 1); z(S_CSTART, NCOL + 2); z(S_META, 8); z(S_CORDER, MAXC); z(S_RK0, MAXC); z(S_RK1, MAXC);
    z(S_RK2, MAXC); z(S_KEPT, MAXC); z(S_OSTART, MAXU); z(S_LEAD, MAXC); z(S_SLOT, (size_t)MAXC * 26); z(S_ROWOF, MAXC); z(S_RUNIT, MAXC);
    z(S_PCOL, MAXC); z(S_PORD, MAXC); z(S_UTC, MAXU); z(S_UTP, MAXU); z(wutp, MAXU); z(S_RCON, MAXC); z(S_UWOFF, UH); z(S_WALK, MAXC);
    z(wkey, MAXU); z(wcol, MAXU); z(wmask, (size_t)std::max(NB, 1u) * NCOLW); z(wcount, NCOL + 2); z(wlinkhit, std::max(NB, 1u));
    z(wjsplit, std::max(NL, 1u));
    z(uf, NB + NA); z(uff, NB + NA);
    tw.assign(std::max(NB, 1u) * 2, V4<T>{0, 0, 0, 0});
    imp.assign(std::max(NB, 1u) * 2, V4<T>{0, 0, 0, 0});
    aqd.assign(std::max(ND, 1u), T(0));
    for (int i = 0; i < 16; i++) stats[i] = 0;
    counters[0] = counters[1] = 0;
  }
};

}  // namespace oracle

// ------------------------------------------------------------------------------------------------
// C entry points (called through ctypes by tests/, smoke() and bench.py's cpu_baseline).
// ------------------------------------------------------------------------------------------------
using namespace oracle;

struct OracleOut {
  // Per-call debug outputs of the last substep (optional; any pointer may be null).
  uint32_t* pairs;      // [2*MAXP]
  uint32_t* npairs;     // counters[0] after broad
  float* contacts;      // [MAXC * 24]: 5 vec4 + (smu0, smu1, 0, 0) in sorted order
  uint32_t* ncontacts;  // counters[1] after narrow
  uint32_t* rowinfo;    // [MAXC * 4]: (contact key lo|hi<<16, color, unit, body pair)
  uint32_t* nrows;
  uint32_t* meta;       // [8]: nu, ncol, nc, ndense, ...
  float* rows;          // [MAXC * ROWS * 4]: contact rows after prepare (initial impulses included)
  // Body-pair reduction inputs and result per sorted contact c (wgsl.js:1459-1503, 2029-2105), or null:
  // red[c*8..]: unit (0xffffffff when the contact has no unit), listing position p, group keys k0, k1, k2,
  // position identity pk, kept (1/0), and the contact's index again; redf[c*4..]: pos.x, pos.y, gap, 0.
  uint32_t* red;        // [MAXC * 8]
  float* redf;          // [MAXC * 4]
};

template <class T>
static int run_substeps(const uint32_t* header, const float* statF, const uint32_t* statU, T* poses, T* vel, T* art, uint32_t* akey,
                        AnchorVal<T>* aval, WarmSlot<T>* wtab, uint32_t* clock, uint32_t nsub, uint32_t* stats_out, OracleOut* dbg) {
  World<T> w;
  w.load_header(header);
  w.statF.resize(header[W_NSTATF]);
  for (uint32_t i = 0; i < header[W_NSTATF]; i++)
    w.statF[i] = {(T)statF[i * 4], (T)statF[i * 4 + 1], (T)statF[i * 4 + 2], (T)statF[i * 4 + 3]};
  w.statU = statU;
  w.poses = (V4<T>*)poses;
  w.vel = (V4<T>*)vel;
  w.art = art;
  w.akey = akey;
  w.aval = aval;
  w.wtab = wtab;
  w.alloc_scratch();
  if (stats_out) for (int i = 0; i < 16; i++) w.stats[i] = stats_out[i];
  // clock: substep count, control flags, frame clock (double, words 2-3)
  uint32_t* substep_count = clock;
  std::memcpy(&w.phase, clock + 2, 8);
  w.control = clock[1];
  for (uint32_t s = 0; s < nsub; s++) {
    w.parity = *substep_count % 2;
    w.substep_in_frame = *substep_count % header[W_SUBSTEPS];
    w.shapes();
    w.broad();
    if (dbg && dbg->pairs) {
      for (size_t i = 0; i < w.pairs.size(); i++) dbg->pairs[i] = w.pairs[i];
      *dbg->npairs = w.counters[0];
    }
    w.narrow();
    if (dbg && dbg->ncontacts) *dbg->ncontacts = w.counters[1];
    w.csort();
    if (dbg && dbg->contacts)
      for (size_t i = 0; i < w.contacts.size(); i++) {
        for (int q = 0; q < 5; q++) {
          dbg->contacts[i * 24 + q * 4 + 0] = (float)w.contacts[i].q[q].x; dbg->contacts[i * 24 + q * 4 + 1] = (float)w.contacts[i].q[q].y;
          dbg->contacts[i * 24 + q * 4 + 2] = (float)w.contacts[i].q[q].z; dbg->contacts[i * 24 + q * 4 + 3] = (float)w.contacts[i].q[q].w;
        }
        dbg->contacts[i * 24 + 20] = (float)w.contacts[i].smu[0]; dbg->contacts[i * 24 + 21] = (float)w.contacts[i].smu[1];
        dbg->contacts[i * 24 + 22] = 0; dbg->contacts[i * 24 + 23] = 0;
      }
    w.prepare();
    if (dbg && dbg->rowinfo) {
      uint32_t nu = w.S_META[0], tot = std::min(w.S_USTART[nu], w.MAXC);
      *dbg->nrows = tot;
      for (uint32_t r = 0; r < tot; r++) {
        uint32_t c = w.S_RCON[r];
        const V4<T>& m = w.contacts[c].q[2];
        dbg->rowinfo[r * 4] = tu(m.z) | (tu(m.w) << 16);
        dbg->rowinfo[r * 4 + 1] = w.S_PCOL[r];
        dbg->rowinfo[r * 4 + 2] = w.S_RUNIT[r];
        {  // the reference's 16-bit packing (a + 1) | (b + 1) << 16 (worlds of fewer than 65536 bodies)
          uint64_t k = w.S_UKEY[w.S_RUNIT[r]];
          dbg->rowinfo[r * 4 + 3] = (uint32_t)(w.key_a(k) + 1) | ((uint32_t)(w.key_b(k) + 1) << 16);
        }
      }
      for (int i = 0; i < 8; i++) dbg->meta[i] = w.S_META[i];
    }
    if (dbg && dbg->red) w.export_reduction(dbg->red, dbg->redf);
    if (dbg && dbg->rows)
      for (size_t i = 0; i < (size_t)w.MAXC * w.ROWS; i++) {
        dbg->rows[i * 4] = (float)w.rows[i].x; dbg->rows[i * 4 + 1] = (float)w.rows[i].y;
        dbg->rows[i * 4 + 2] = (float)w.rows[i].z; dbg->rows[i * 4 + 3] = (float)w.rows[i].w;
      }
    w.solve();
    *substep_count += 1;
    if (*substep_count % header[W_SUBSTEPS] == 0) w.phase += 1.0 / (double)u2f(header[W_FPS]);
  }
  std::memcpy(clock + 2, &w.phase, 8);
  if (stats_out) for (int i = 0; i < 16; i++) stats_out[i] = w.stats[i];
  return 0;
}

extern "C" {

// Size in bytes of one anchor-table value record and of one warm-start table slot for the float and double oracles.
size_t oracle_anchor_val_size_f32() { return sizeof(AnchorVal<float>); }
size_t oracle_anchor_val_size_f64() { return sizeof(AnchorVal<double>); }
size_t oracle_warm_slot_size_f32() { return sizeof(WarmSlot<float>); }
size_t oracle_warm_slot_size_f64() { return sizeof(WarmSlot<double>); }

// Runs nsub substeps on one world, in place.  poses/vel/art are float; akey is uint32[2*AH]; aval is
// opaque storage of 2*AH * oracle_anchor_val_size_f32() bytes (zeroed at reset); warm is opaque storage of
// 2*W_WH * oracle_warm_slot_size_f32() bytes (zeroed at reset; unused when W_WH is 0); clock is uint32[4]: substep
// count, control flags (bit 0: user-set targets), frame clock in seconds (double), all zero at reset.
int oracle_substeps_f32(const uint32_t* header, const float* statF, const uint32_t* statU, float* poses, float* vel, float* art,
                        uint32_t* akey, void* aval, void* warm, uint32_t* clock, uint32_t nsub, uint32_t* stats, OracleOut* dbg) {
  return run_substeps<float>(header, statF, statU, poses, vel, art, akey, (AnchorVal<float>*)aval, (WarmSlot<float>*)warm, clock, nsub,
                             stats, dbg);
}
int oracle_substeps_f64(const uint32_t* header, const float* statF, const uint32_t* statU, double* poses, double* vel, double* art,
                        uint32_t* akey, void* aval, void* warm, uint32_t* clock, uint32_t nsub, uint32_t* stats, OracleOut* dbg) {
  return run_substeps<double>(header, statF, statU, poses, vel, art, akey, (AnchorVal<double>*)aval, (WarmSlot<double>*)warm, clock,
                              nsub, stats, dbg);
}
// Anchor table of the float oracle in the kernels' uint32 layout (keys [2*AH], then AVN words per slot):
// lets a test compare the carried anchors entry by entry.
void oracle_anchor_export_f32(uint32_t AH, const uint32_t* akey, const void* aval, uint32_t* out) {
  const AnchorVal<float>* v = (const AnchorVal<float>*)aval;
  for (uint32_t i = 0; i < 2 * AH; i++) out[i] = akey[i];
  for (uint32_t i = 0; i < 2 * AH; i++) {
    uint32_t* o = out + 2 * AH + i * AVN;
    o[0] = v[i].pair; o[1] = v[i].k;
    o[2] = f2u(v[i].ll.x); o[3] = f2u(v[i].ll.y); o[4] = f2u(v[i].ll.z);
    o[5] = f2u(v[i].el.x); o[6] = f2u(v[i].el.y); o[7] = f2u(v[i].el.z);
    o[8] = f2u(v[i].il.x); o[9] = f2u(v[i].il.y); o[10] = f2u(v[i].il.z);
    o[11] = f2u(v[i].nl.x); o[12] = f2u(v[i].nl.y); o[13] = f2u(v[i].nl.z);
  }
}
// Warm-start table of the float oracle in the kernels' uint32 layout (keys [2*WH], then 4 words per slot: normal
// impulse, tangent impulse as a world vector).
void oracle_warm_export_f32(uint32_t WH, const void* warm, uint32_t* out) {
  const WarmSlot<float>* s = (const WarmSlot<float>*)warm;
  for (uint32_t i = 0; i < 2 * WH; i++) {
    out[i] = s[i].key;
    uint32_t* o = out + 2 * WH + i * 4;
    o[0] = f2u(s[i].v.x); o[1] = f2u(s[i].v.y); o[2] = f2u(s[i].v.z); o[3] = f2u(s[i].v.w);
  }
}
// One world's float-oracle buffers (as oracle_substeps_f32 takes them), for the threaded batch runner.
struct OracleWorldPtrs {
  const uint32_t* header;
  const float* statF;
  const uint32_t* statU;
  float* poses;
  float* vel;
  float* art;
  uint32_t* akey;
  void* aval;
  void* warm;
  uint32_t* clock;
  uint32_t* stats;
};

// CPU baseline runner: n worlds on `nthreads` std::threads (worlds dealt from a shared counter).  Every world
// first steps `frames_untimed` frames; after all threads have finished that phase, every world steps
// `frames_timed` frames, and timed_wall_s receives the wall time of that second phase (first world start to
// last world end).  busy_s[t] receives thread t's stepping time in the timed phase.
int oracle_run_batch_f32(uint32_t n, const OracleWorldPtrs* w, uint32_t frames_untimed, uint32_t frames_timed, uint32_t nthreads,
                         double* timed_wall_s, double* busy_s) {
  if (nthreads == 0) nthreads = 1;
  auto phase = [&](uint32_t frames, bool timed) {
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < nthreads; t++)
      th.emplace_back([&, t]() {
        double busy = 0;
        for (;;) {
          uint32_t i = next.fetch_add(1);
          if (i >= n) break;
          const OracleWorldPtrs& x = w[i];
          auto t0 = std::chrono::steady_clock::now();
          run_substeps<float>(x.header, x.statF, x.statU, x.poses, x.vel, x.art, x.akey, (AnchorVal<float>*)x.aval, (WarmSlot<float>*)x.warm,
                              x.clock, frames * x.header[W_SUBSTEPS], x.stats, nullptr);
          busy += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        if (timed && busy_s) busy_s[t] = busy;
      });
    for (auto& x : th) x.join();
  };
  if (frames_untimed) phase(frames_untimed, false);
  auto t0 = std::chrono::steady_clock::now();
  if (frames_timed) phase(frames_timed, true);
  if (timed_wall_s) *timed_wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// One-step error reference: from the float oracle state of a world (poses, vel, art, anchors, warm table, clock),
// steps `nsub` substeps in float in place and, from the same state widened to double, in double; the double
// results go to poses64 / vel64 / art64.  wall_s[0] / wall_s[1] receive the float and double stepping times.
int oracle_step_pair_f32_f64(const OracleWorldPtrs* w, uint32_t nsub, double* poses64, double* vel64, double* art64, double* wall_s) {
  const uint32_t* h = w->header;
  size_t np = (size_t)h[W_NPOSES] * 4, nv = (size_t)std::max(h[W_NB], 1u) * 8, na = h[W_NART];
  for (size_t i = 0; i < np; i++) poses64[i] = w->poses[i];
  for (size_t i = 0; i < nv; i++) vel64[i] = w->vel[i];
  for (size_t i = 0; i < na; i++) art64[i] = w->art[i];
  uint32_t AH = h[W_AH], WH = h[W_WH];
  std::vector<uint32_t> akey(w->akey, w->akey + 2 * AH);
  std::vector<AnchorVal<double>> aval(2 * AH);
  const AnchorVal<float>* af = (const AnchorVal<float>*)w->aval;
  auto v3 = [](V3<float> x) { return V3<double>{x.x, x.y, x.z}; };
  for (uint32_t i = 0; i < 2 * AH; i++) aval[i] = {af[i].pair, af[i].k, v3(af[i].ll), v3(af[i].el), v3(af[i].il), v3(af[i].nl)};
  std::vector<WarmSlot<double>> warm(std::max(2 * WH, 1u));
  const WarmSlot<float>* wf = (const WarmSlot<float>*)w->warm;
  for (uint32_t i = 0; i < 2 * WH; i++) warm[i] = {wf[i].key, V4<double>{wf[i].v.x, wf[i].v.y, wf[i].v.z, wf[i].v.w}};
  uint32_t clock64[4], stats64[16];
  for (int i = 0; i < 4; i++) clock64[i] = w->clock[i];
  for (int i = 0; i < 16; i++) stats64[i] = w->stats[i];
  auto t0 = std::chrono::steady_clock::now();
  run_substeps<float>(h, w->statF, w->statU, w->poses, w->vel, w->art, w->akey, (AnchorVal<float>*)w->aval, (WarmSlot<float>*)w->warm,
                      w->clock, nsub, w->stats, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  run_substeps<double>(h, w->statF, w->statU, poses64, vel64, art64, akey.data(), aval.data(), warm.data(), clock64, nsub, stats64, nullptr);
  auto t2 = std::chrono::steady_clock::now();
  if (wall_s) {
    wall_s[0] = std::chrono::duration<double>(t1 - t0).count();
    wall_s[1] = std::chrono::duration<double>(t2 - t1).count();
  }
  return 0;
}

// Known-answer hooks for the oracle's own pinning tests.
// World of static data only (geometry queries): header, statF and statU of a packed world.
static World<double> geometry_world(const uint32_t* header, const float* statF, const uint32_t* statU) {
  World<double> w;
  w.load_header(header);
  w.statF.resize(header[W_NSTATF]);
  for (uint32_t i = 0; i < header[W_NSTATF]; i++)
    w.statF[i] = {(double)statF[i * 4], (double)statF[i * 4 + 1], (double)statF[i * 4 + 2], (double)statF[i * 4 + 3]};
  w.statU = statU;
  return w;
}
// MPR of hulls ha (frame of reference) and hb (rotation rq = x, y, z, w and offset t in ha's frame), shapes
// inflated by r: out = (normal, depth), depth < 0 when they miss.
void oracle_mpr_f64(const uint32_t* header, const float* statF, const uint32_t* statU, uint32_t ha, uint32_t hb, const double* rq,
                    const double* t, double r, double* out) {
  World<double> w = geometry_world(header, statF, statU);
  V4<double> m = w.mpr(ha, hb, V4<double>{rq[0], rq[1], rq[2], rq[3]}, V3<double>{t[0], t[1], t[2]}, r);
  out[0] = m.x; out[1] = m.y; out[2] = m.z; out[3] = m.w;
}
// Support point of shape s's collision geometry along d (shape frame).
void oracle_gsup_f64(const uint32_t* header, const float* statF, const uint32_t* statU, uint32_t s, const double* d, double* out) {
  World<double> w = geometry_world(header, statF, statU);
  V3<double> p = w.gsup(w.geo(s), V3<double>{d[0], d[1], d[2]});
  out[0] = p.x; out[1] = p.y; out[2] = p.z;
}
// Featherstone probes on a world's current state, for the pins against an independent composite-rigid-body
// mass matrix.  Runs FK and the articulated-body factor of every articulation as prepare does
// (artwgsl.js:98-369), then:
//   mode 0: link a0's unit-impulse response at its COM (artwgsl.js:431-510): out[0..35] = W2 (unsplit),
//           out[36..71] = W with the path's split counts in[0..path length) (in = null: all 1);
//   mode 1: dof_hinv(link a0, its dof a1) (artwgsl.js:513-518): out[0];
//   mode 2: art_solve_gen(articulation a0, tau = in[0..NDA)) (artwgsl.js:557-596): out[0..NDA);
//   mode 3: add_gen_jac(link a0, point in[0..2], directions in[3..5], in[6..8], in[9..11]) (artwgsl.js:525-540),
//           relative to the articulation's first dof: out[0..3*NDA), zeroed first;
//   mode 4: every link's frame after FK: out[l*10..] = position (3), rotation quaternion xyzw (4), COM (3).
// poses, vel and art are doubles and are not written back except through art (FK and factor outputs).
int oracle_art_probe_f64(const uint32_t* header, const float* statF, const uint32_t* statU, double* poses, double* vel, double* art,
                         uint32_t mode, uint32_t a0, uint32_t a1, const double* in, double* out) {
  World<double> w;
  w.load_header(header);
  w.statF.resize(header[W_NSTATF]);
  for (uint32_t i = 0; i < header[W_NSTATF]; i++)
    w.statF[i] = {(double)statF[i * 4], (double)statF[i * 4 + 1], (double)statF[i * 4 + 2], (double)statF[i * 4 + 3]};
  w.statU = statU;
  w.poses = (V4<double>*)poses;
  w.vel = (V4<double>*)vel;
  w.art = art;
  w.alloc_scratch();
  for (uint32_t a = 0; a < w.NA; a++) { w.art_fk(a); w.art_factor(a); }
  if (mode == 0) {
    if (a0 >= w.NL) return -1;
    std::vector<uint32_t> path(w.MAXPATH), cnt(w.MAXPATH, 1u);
    uint32_t np = w.link_path(a0, path.data());
    if (in) for (uint32_t i = 0; i < np; i++) cnt[i] = (uint32_t)in[i];
    w.link_response(a0, cnt.data());
    for (uint32_t k = 0; k < 36; k++) { out[k] = art[w.lo(a0) + L_W2 + k]; out[36 + k] = art[w.lo(a0) + L_W + k]; }
  } else if (mode == 1) {
    if (a0 >= w.NL || a1 >= w.link_ndof(a0)) return -1;
    out[0] = w.dof_hinv(a0, a1);
  } else if (mode == 2) {
    if (a0 >= w.NA) return -1;
    std::vector<double> y(w.NDA, 0.0);
    w.art_solve_gen(a0, in, y.data());
    for (uint32_t k = 0; k < w.NDA; k++) out[k] = y[k];
  } else if (mode == 3) {
    if (a0 >= w.NL) return -1;
    int b = w.link_body(a0);
    uint32_t ai[4];
    w.art_info((uint32_t)w.body_art(b), ai);
    for (uint32_t k = 0; k < 3 * w.NDA; k++) out[k] = 0;
    w.add_gen_jac(a0, V3<double>{in[0], in[1], in[2]}, 1.0, V3<double>{in[3], in[4], in[5]}, V3<double>{in[6], in[7], in[8]},
                  V3<double>{in[9], in[10], in[11]}, ai[2], out);
  } else if (mode == 4) {
    for (uint32_t l = 0; l < w.NL; l++) {
      V3<double> p = w.ld3(w.lo(l) + L_POS), c = w.ld3(w.lo(l) + L_COM);
      V4<double> q = w.link_rot(l);
      double* o = out + l * 10;
      o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = q.x; o[4] = q.y; o[5] = q.z; o[6] = q.w; o[7] = c.x; o[8] = c.y; o[9] = c.z;
    }
  } else {
    return -1;
  }
  return 0;
}
uint32_t oracle_hash_u32(uint32_t x) { return hash_u32(x); }
float oracle_sinf(float x) { return sinf_c(x); }
float oracle_cosf(float x) { return cosf_c(x); }
float oracle_atan2f(float y, float x) { return atan2f_c(y, x); }
void oracle_friction_pair_solve(float a, float c, float d, float r0, float r1, float o0, float o1, float radius, float omega, float* out) {
  V2<float> r = World<float>::friction_pair_solve(a, c, d, r0, r1, o0, o1, radius, omega);
  out[0] = r.x; out[1] = r.y;
}
// The articulated spin/rolling cone block in FP64 (relaxation omega, creep speed creep): H0[25], g0[5], mu[5], lam[5],
// load; out[5].
void oracle_ell_block_f64(const double* H0, const double* g0, const double* mu, const double* lam, double load, double omega,
                          double creep, double* out) {
  World<double> w;
  w.F_OMEGA = omega;
  w.F_CREEP = creep;
  w.ell_block(H0, g0, mu, lam, load, out);
}
// The coupled contact block in FP64: in[0] un, in[1..2] ut0, in[3..5] (gnn, gnt), in[6..8] (g_t1t1, g_t1t2,
// g_t2t2), in[9] vtarget, in[10] mu, in[11..13] lam; out (ln, lt1, lt2).
void oracle_contact_block_f64(uint32_t mode, const double* in, double* out) {
  V3<double> r = World<double>::contact_block(mode, in[0], V2<double>{in[1], in[2]}, V4<double>{in[3], in[4], in[5], 0},
                                              V4<double>{in[6], in[7], in[8], 0}, in[9], in[10], V3<double>{in[11], in[12], in[13]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
}
